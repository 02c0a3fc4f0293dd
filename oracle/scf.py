"""FP64 CPU restatement of SPARC's density-mixing SCF loop, built from the
oracle's C primitives (filter, projection, generalized eigensolve, rotation,
Lanczos, AAR, Fermi level, density, LDA-PZ).

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of the device SCF.

Steps, in the reference's order (src/electronicGroundState.c:749-1006):
  Calculate_elecDens        CheFSI pass(es), occupations, rho_out
  Calculate_Free_Energy     energies evaluated with rho_in (src/energy.c:41-199)
  Evaluate_scf_error        |rho_out - rho_in| / |rho_out| (src/electronicGroundState.c:1160)
  Mixing_periodic_pulay     Pulay + Kerker density mixing (src/mixing.c:213-527)
  Calculate_elecstPotential / Calculate_Vxc / Calculate_Veff_loc_dmcomm_phi
"""
from __future__ import annotations

import math

import numpy as np

from .oracle import Oracle, anderson_coeff, density, fermi_level, lda, smearing, sygv


def entropy(lam, occ, ef, beta, smear_type, occfac):
    """Calculate_electronicEntropy / Calculate_entropy_term (src/energy.c:241-330)."""
    s = 0.0
    if smear_type == 0:
        for g in occ:
            if g > 1e-12 and 1.0 - g > 1e-12:
                s += -(g * math.log(g) + (1.0 - g) * math.log(1.0 - g))
    else:
        c = 0.5 / math.sqrt(math.pi)
        for l in lam:
            x = beta * (l - ef)
            s += c * math.exp(-x * x)
    return -occfac / beta * s


class OracleSCF:
    """State of one SCF run; ``run()`` iterates to convergence."""

    def __init__(self, grid, nloc, b, rho0, params: dict, psi0: np.ndarray, x0: np.ndarray,
                 rho_core=None):
        self.g = grid
        self.rho_core = None if rho_core is None else np.ascontiguousarray(rho_core, np.float64).ravel()
        self.p = dict(params)
        self.o = Oracle(grid, None, nloc)
        nd = grid.nd
        self.b = np.ascontiguousarray(b, np.float64).ravel()
        self.rho_in = np.array(rho0, np.float64).ravel()
        self.x_km1 = self.rho_in.copy()
        self.phi = np.zeros(nd)
        self.f_k = np.zeros(nd)
        self.f_km1 = np.zeros(nd)
        m = self.p["mix_history"]
        self.R = np.zeros((nd, m), order="F")
        self.F = np.zeros((nd, m), order="F")
        self.Pf = np.zeros(nd)
        self.X = np.asfortranarray(np.array(psi0, np.float64))
        self.x0 = np.ascontiguousarray(x0, np.float64).ravel()
        self.lam = None
        self.occ = None
        self.ef = 0.0
        self.eigmin = self.eigmax = 0.0
        self.history = []
        self.poisson_iters = 0
        self.kerker_iters = 0
        self.potential(self.rho_in)

    def potential(self, rho):
        """phi = Poisson(4 pi (b + rho)) warm-started, mean zero
        (src/electrostatics.c:1593-1679); Vxc = LDA-PZ; Veff = phi + Vxc."""
        p = self.p
        rhs = 4.0 * math.pi * (self.b + rho)
        self.phi, it, _ = self.o.poisson_solve(self.phi, rhs, p["tol_poisson"], p["maxit_poisson"])
        self.poisson_iters += it
        # add_rho_core: XC evaluated at max(rho + rho_core, xc_rhotol = 1e-14)
        # (src/exchangeCorrelation.c:1254-1270, src/initialization.c:1843)
        rxc = np.maximum(rho if self.rho_core is None else rho + self.rho_core, 1e-14)
        exc, self.vxc = lda(rxc, p.get("xc", 0))
        self.exc_density = rxc * exc
        self.o.veff = self.phi + self.vxc

    def chefsi(self, count: int):
        """eigSolve_CheFSI + Chebyshevfilter_constants (src/eigenSolver.c:150-263,589-715)."""
        p = self.p
        if count == 0:
            emin, emax, _ = self.o.lanczos(self.x0, p["tol_lanczos"], p["tol_lanczos"], 1000)
            self.eigmax = emax * 1.01
            self.eigmin = emin - 0.1
            cutoff = 0.5 * (self.eigmin + self.eigmax)
        else:
            if count >= p["rho_trigger"]:
                self.eigmin = self.lam[0]
            cutoff = self.lam[-1] + 0.1
        Y, _ = self.o.chebyshev_filter(self.X, p["cheb_degree"], cutoff, self.eigmax, self.eigmin)
        Hp, Mp, _ = self.o.project(Y)
        self.lam, Q = sygv(Hp, Mp)
        self.X = self.o.rotate(Y, Q)
        # Calculate_occupation (src/occupation.c:43-123): bracket [lambda_0 - 1, lambda_end + 1]
        ef, st = fermi_level(self.lam, p["occfac"], p["nelec"], p["beta"], p["smear_type"],
                             self.lam[0] - 1.0, self.lam[-1] + 1.0, 1e-12, 100)
        if st != 0:
            raise RuntimeError("oracle SCF: Fermi level not bracketed")
        self.ef = ef
        self.occ = smearing(p["beta"], self.lam, ef, p["smear_type"])

    def step(self, scf: int) -> dict:
        p = self.p
        g = self.g
        dV = g.dV
        npass = p["rho_trigger"] if scf == 0 else 1
        veff_used = self.o.veff.copy()
        for k in range(npass):
            self.chefsi(k if scf == 0 else p["rho_trigger"] + scf - 1)
        rho_out = density(self.X, self.occ, p["occfac"], dV)
        rho_in = self.rho_in
        eband = float(np.sum(p["occfac"] * self.occ * self.lam))
        exc = float(self.exc_density.sum()) * dV
        e1 = 0.5 * float(self.b @ self.phi) * dV
        e2 = 0.5 * float(rho_in @ self.phi) * dV
        e3 = float(rho_in @ self.vxc) * dV
        ent = entropy(self.lam, self.occ, self.ef, p["beta"], p["smear_type"], p["occfac"])
        etot = eband + e1 - e2 - e3 + exc + p["Esc"] + ent
        d = rho_out - rho_in
        err = math.sqrt(float(d @ d) / float(rho_out @ rho_out))
        rec = dict(Etot=etot, Eband=eband, Exc=exc, E1=e1, E2=e2, E3=e3, Entropy=ent, Efermi=self.ef,
                   scf_error=err, lambda_=self.lam.copy(), veff=veff_used, rho_in=rho_in,
                   rho_out=rho_out, phi=self.phi.copy(), vxc=self.vxc.copy())
        self.mix(scf, rho_out)
        self.potential(self.rho_in)
        return rec

    def mix(self, scf: int, g_k: np.ndarray):
        """Mixing_periodic_pulay + Kerker_precond (src/mixing.c:213-527), density mixing."""
        p = self.p
        m = p["mix_history"]
        pulay = ((scf + 1) % p["pulay_freq"] == 0) and scf > 0
        amix = p["mix_beta"] if pulay else p["mix_omega"]
        x_k = self.rho_in
        if scf > 0:
            self.f_km1 = self.f_k
        self.f_k = g_k - x_k
        if scf > 0:
            ih = (scf - 1) % m
            self.R[:, ih] = x_k - self.x_km1
            self.F[:, ih] = self.f_k - self.f_km1
        if pulay:
            G = anderson_coeff(self.f_k, self.F)
            x_w = x_k - self.R @ G
            f_w = self.f_k - self.F @ G
        else:
            x_w, f_w = x_k.copy(), self.f_k.copy()
        k2 = p["kerker_ktf"] ** 2
        Lf = self.o.lap_plus_diag(f_w, 1.0, 0.0, -k2 * p["kerker_thresh"]).ravel()
        self.Pf, it, _ = self.o.aar(self.Pf, Lf, c=-k2, tol=p["tol_precond"], max_iter=1000)
        self.kerker_iters += it
        if abs(p["kerker_ktf"]) < 1e-14:
            self.Pf -= self.Pf.sum() / self.g.nd
        self.Pf *= -amix
        x_kp1 = np.maximum(x_w + self.Pf, 0.0)
        x_kp1 *= p["nelec"] / (x_kp1.sum() * self.g.dV)
        self.x_km1 = x_k
        self.rho_in = x_kp1

    def run(self) -> dict:
        p = self.p
        conv = False
        for scf in range(p["max_scf"]):
            rec = self.step(scf)
            self.history.append(rec)
            if rec["scf_error"] < p["tol_scf"] and scf + 1 >= p["min_scf"]:
                conv = True
                break
        out = dict(self.history[-1])
        out.update(scf_iters=len(self.history), converged=conv, rho=self.rho_in.copy(),
                   phi=self.phi.copy(), X=self.X, iter_err=np.array([h["scf_error"] for h in self.history]),
                   iter_etot=np.array([h["Etot"] for h in self.history]),
                   poisson_iters=self.poisson_iters, kerker_iters=self.kerker_iters)
        return out
