"""Grid, nonlocal-projector containers and the synthetic operator workloads.

The containers mirror the fields SPARC's hot path reads (orthogonal grid,
FD order, boundary conditions; ATOM_NLOC_INFLUENCE_OBJ / NLOC_PROJ_OBJ,
src/include/isddft.h:202-216,247-265) in flat numpy form, which is the form the
C ABI of libsparcgpu.so takes.

Workloads:
  * operator_filter_workload: the 96^3 x 512 Chebyshev-filter operator case
    (Gaussian Veff, 64 Gaussian x real-Ylm projector sites of 8 projectors,
    6 Bohr cutoff, degree 24, bounds from a 10-step Lanczos).
  * poisson_workload: 128^3 periodic Poisson right-hand side from 216 Gaussian
    pseudocharges of charge +4 on a neutralising background.
Both accept smaller sizes for parity tests.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# ---------------------------------------------------------------- grid


@dataclass
class Grid:
    nx: int
    ny: int
    nz: int
    hx: float
    hy: float
    hz: float
    bcx: int = 0  # 0 periodic, 1 Dirichlet (pSPARC->BCx convention)
    bcy: int = 0
    bcz: int = 0
    order: int = 12

    @property
    def nd(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def dV(self) -> float:
        return self.hx * self.hy * self.hz

    @property
    def lengths(self) -> tuple[float, float, float]:
        # periodic: L = N h; Dirichlet: L = (N + 1) h  (src/initialization.c:1936 mesh rule)
        return tuple((n + (1 if bc else 0)) * h for n, h, bc in
                     ((self.nx, self.hx, self.bcx), (self.ny, self.hy, self.bcy),
                      (self.nz, self.hz, self.bcz)))

    def coords(self):
        """x, y, z coordinates (flattened, x fastest) of every grid point."""
        x = np.arange(self.nx) * self.hx
        y = np.arange(self.ny) * self.hy
        z = np.arange(self.nz) * self.hz
        Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
        return X.ravel(), Y.ravel(), Z.ravel()


# ---------------------------------------------------------------- nonlocal


@dataclass
class NonlocalSet:
    """Flattened projector set: one entry per influencing (image) atom."""
    ndc: np.ndarray        # int32 [n_influence]
    nproj: np.ndarray      # int32 [n_influence]
    pos_off: np.ndarray    # int64 [n_influence]
    grid_pos: np.ndarray   # int32 [sum ndc]
    chi_off: np.ndarray    # int64 [n_influence]
    chi: np.ndarray        # float64 [sum ndc*nproj], column-major ndc x nproj per entry
    alpha_off: np.ndarray  # int32 [n_influence]
    n_alpha: int
    gamma: np.ndarray      # float64 [n_alpha]
    dV: float

    @property
    def n_influence(self) -> int:
        return int(self.ndc.shape[0])

    def dense(self, nd: int) -> np.ndarray:
        """Dense Vnl = sum_J Chi_J Gamma_J Chi_J' dV (only for tiny grids)."""
        # images share alpha rows, so build the per-row projector vectors first
        P = np.zeros((nd, self.n_alpha))
        for a in range(self.n_influence):
            n, p = int(self.ndc[a]), int(self.nproj[a])
            pos = self.grid_pos[self.pos_off[a]:self.pos_off[a] + n]
            C = self.chi[self.chi_off[a]:self.chi_off[a] + n * p].reshape(p, n).T
            for k in range(p):
                np.add.at(P[:, self.alpha_off[a] + k], pos, C[:, k])
        return (P * self.gamma) @ P.T * self.dV


# C(l,m) prefactors of the real spherical harmonics (SPARC's conventions,
# src/tools.c:1231-1330): l=1 uses y, z, x for m = -1, 0, 1.
def real_ylm(l: int, m: int, x, y, z, r):
    with np.errstate(invalid="ignore", divide="ignore"):
        if l == 0:
            return np.full_like(r, 0.5 * np.sqrt(1.0 / np.pi))
        rr = np.where(r > 0, r, 1.0)
        if l == 1:
            c = np.sqrt(3.0 / (4.0 * np.pi))
            v = {-1: y, 0: z, 1: x}[m] / rr
            return np.where(r > 0, c * v, 0.0)
        if l == 2:
            r2 = rr * rr
            if m == -2:
                v = 0.5 * np.sqrt(15.0 / np.pi) * x * y / r2
            elif m == -1:
                v = 0.5 * np.sqrt(15.0 / np.pi) * y * z / r2
            elif m == 0:
                v = 0.25 * np.sqrt(5.0 / np.pi) * (-x * x - y * y + 2 * z * z) / r2
            elif m == 1:
                v = 0.5 * np.sqrt(15.0 / np.pi) * z * x / r2
            else:
                v = 0.25 * np.sqrt(15.0 / np.pi) * (x * x - y * y) / r2
            return np.where(r > 0, v, 0.0)
    raise ValueError("l > 2 not supported")


def _periodic_delta(d, L, periodic):
    return d - L * np.round(d / L) if periodic else d


def sphere_points(grid: Grid, center, rc: float):
    """Grid points within rc of center, minimum image for periodic axes.

    Returns (linear indices sorted ascending, dx, dy, dz) -- displacement from
    the center.  Requires rc < L/2 on periodic axes (no self-overlap)."""
    L = grid.lengths
    out = []
    for ax, (n, h, bc, c, Lx) in enumerate(((grid.nx, grid.hx, grid.bcx, center[0], L[0]),
                                            (grid.ny, grid.hy, grid.bcy, center[1], L[1]),
                                            (grid.nz, grid.hz, grid.bcz, center[2], L[2]))):
        if not bc and 2 * rc >= Lx:
            raise ValueError("projector cutoff must be below half the periodic cell")
        lo = int(np.floor((c - rc) / h)) - 1
        hi = int(np.ceil((c + rc) / h)) + 1
        idx = np.arange(lo, hi + 1)
        if bc:
            idx = idx[(idx >= 0) & (idx < n)]
            d = idx * h - c
        else:
            d = idx * h - c
            idx = np.mod(idx, n)
        out.append((idx, d))
    (ix, dx), (iy, dy), (iz, dz) = out
    DZ, DY, DX = np.meshgrid(dz, dy, dx, indexing="ij")
    IZ, IY, IX = np.meshgrid(iz, iy, ix, indexing="ij")
    r2 = DX ** 2 + DY ** 2 + DZ ** 2
    mask = r2 <= rc * rc
    lin = (IZ[mask].astype(np.int64) * grid.ny + IY[mask]) * grid.nx + IX[mask]
    dxs, dys, dzs = DX[mask], DY[mask], DZ[mask]
    order = np.argsort(lin, kind="stable")
    lin = lin[order]
    if np.unique(lin).size != lin.size:
        raise ValueError("projector sphere overlaps its own periodic image")
    return lin, dxs[order], dys[order], dzs[order]


def gaussian_projectors(grid: Grid, sites: np.ndarray, rc: float, sigmas, gammas, lmax: int = 1):
    """Projector set with Chi = exp(-r^2 / (2 sigma_p^2)) Y_lm per site.

    Projector order per site follows SPARC (src/nlocVecRoutines.c:841-863):
    l ascending, then radial index p, then m = -l..l."""
    ndc, nproj, pos_off, chi_off, alpha_off = [], [], [], [], []
    pos_all, chi_all, gam_all = [], [], []
    po = co = ao = 0
    for s in sites:
        lin, dx, dy, dz = sphere_points(grid, s, rc)
        r = np.sqrt(dx * dx + dy * dy + dz * dz)
        cols, g = [], []
        for l in range(lmax + 1):
            for p, (sg, gm) in enumerate(zip(sigmas, gammas)):
                rad = np.exp(-r * r / (2.0 * sg * sg))
                for m in range(-l, l + 1):
                    cols.append(rad * real_ylm(l, m, dx, dy, dz, r))
                    g.append(gm)
        C = np.stack(cols, axis=1)  # ndc x nproj
        ndc.append(lin.size); nproj.append(C.shape[1])
        pos_off.append(po); chi_off.append(co); alpha_off.append(ao)
        pos_all.append(lin.astype(np.int32))
        chi_all.append(C.T.ravel())  # column-major ndc x nproj
        gam_all.extend(g)
        po += lin.size; co += C.size; ao += C.shape[1]
    return NonlocalSet(
        ndc=np.asarray(ndc, np.int32), nproj=np.asarray(nproj, np.int32),
        pos_off=np.asarray(pos_off, np.int64), grid_pos=np.concatenate(pos_all),
        chi_off=np.asarray(chi_off, np.int64), chi=np.concatenate(chi_all),
        alpha_off=np.asarray(alpha_off, np.int32), n_alpha=ao,
        gamma=np.asarray(gam_all, np.float64), dV=grid.dV)


def gaussian_veff(grid: Grid, n: int, rng: np.random.Generator, amp=(-2.0, 0.0), width=(1.0, 3.0)):
    """Veff = sum_i A_i exp(-|r - R_i|^2 / (2 w_i^2)), minimum-image distances."""
    X, Y, Z = grid.coords()
    L = grid.lengths
    A = rng.uniform(*amp, size=n)
    W = rng.uniform(*width, size=n)
    R = rng.uniform(0, 1, size=(n, 3)) * np.asarray(L)
    v = np.zeros(grid.nd)
    for a, w, c in zip(A, W, R):
        dx = _periodic_delta(X - c[0], L[0], not grid.bcx)
        dy = _periodic_delta(Y - c[1], L[1], not grid.bcy)
        dz = _periodic_delta(Z - c[2], L[2], not grid.bcz)
        v += a * np.exp(-(dx * dx + dy * dy + dz * dz) / (2.0 * w * w))
    return v


# ---------------------------------------------------------------- RNG

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _mix64(z: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (z + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        return z ^ (z >> np.uint64(31))


def uniform_block(nd: int, ncol: int, seed: int, lo: float, hi: float) -> np.ndarray:
    """Host copy of sgpu_block_fill_uniform: element (row, col) of an nd x ncol
    column-major block from the counter-based hash of (seed, col*nd + row)."""
    i = np.arange(nd * ncol, dtype=np.uint64)
    r = _mix64(np.uint64(seed) ^ _mix64(i))
    u = (r >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return (lo + (hi - lo) * u).reshape(ncol, nd).T.copy(order="F")


# ---------------------------------------------------------------- workloads


@dataclass
class FilterWorkload:
    grid: Grid
    veff: np.ndarray
    nloc: NonlocalSet
    ncol: int
    degree: int
    psi_seed: int
    lanczos_seed: int
    meta: dict = field(default_factory=dict)


def operator_filter_workload(n: int = 96, h: float = 0.3, ncol: int = 512, n_sites: int = 64,
                             rc: float = 6.0, degree: int = 24, n_gauss: int = 16,
                             seed: int = 2025) -> FilterWorkload:
    """The operator-level filter case: n^3 periodic grid of mesh h, Veff from
    n_gauss Gaussians (A in [-2,0] Ha, w in [1,3] Bohr), n_sites projector
    sites x 8 projectors (2 radial x l<=1), cutoff rc, degree `degree`."""
    g = Grid(n, n, n, h, h, h)
    rng = np.random.default_rng(seed)
    veff = gaussian_veff(g, n_gauss, rng)
    sites = rng.uniform(0, 1, size=(n_sites, 3)) * np.asarray(g.lengths)
    rc_eff = min(rc, 0.49 * min(g.lengths))
    nl = gaussian_projectors(g, sites, rc_eff, sigmas=(1.5, 1.0), gammas=(1.0, -1.0), lmax=1)
    return FilterWorkload(g, veff, nl, ncol, degree, psi_seed=seed + 1, lanczos_seed=seed + 2,
                          meta={"n": n, "h": h, "ncol": ncol, "n_sites": n_sites, "rc": rc_eff,
                                "degree": degree})


def poisson_workload(n: int = 128, h: float = 0.35, n_charges: int = 216, q: float = 4.0,
                     width=(0.8, 1.5), seed: int = 77):
    """rhs = 4 pi (rho - mean rho) for n_charges normalised Gaussian charges."""
    g = Grid(n, n, n, h, h, h)
    rng = np.random.default_rng(seed)
    X, Y, Z = g.coords()
    L = g.lengths
    rho = np.zeros(g.nd)
    W = rng.uniform(*width, size=n_charges)
    R = rng.uniform(0, 1, size=(n_charges, 3)) * np.asarray(L)
    for w, c in zip(W, R):
        cut = 6.0 * w
        dx = _periodic_delta(X - c[0], L[0], True)
        dy = _periodic_delta(Y - c[1], L[1], True)
        dz = _periodic_delta(Z - c[2], L[2], True)
        r2 = dx * dx + dy * dy + dz * dz
        m = r2 < cut * cut
        rho[m] += q * np.exp(-r2[m] / (2 * w * w)) / ((2 * np.pi) ** 1.5 * w ** 3)
    rho -= rho.mean()
    return g, 4.0 * np.pi * rho
