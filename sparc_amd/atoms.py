"""Atom setup from psp8 pseudopotentials on an orthogonal periodic grid:
pseudocharge cutoff radii, pseudocharge density b, the self + correction
energy Esc, the superposed isolated-atom guess density, the NLCC core
density, and the Kleinman-Bylander projector set.

Follows SPARC (orthogonal cells, cell_typ 0, one domain owning the grid):
  * Calculate_PseudochargeCutoff   src/electrostatics.c:44-577
  * GetInfluencingAtoms_local       src/electrostatics.c:585-770
  * Generate_PseudoChargeDensity   src/electrostatics.c:772-1255
  * Calculate_Pseudopot_Ref        src/electrostatics.c:1465-1483
  * GetInfluencingAtoms_nloc       src/nlocVecRoutines.c:43-386
  * CalculateNonlocalProjectors    src/nlocVecRoutines.c:394-571
  * RealSphericalHarmonic          src/tools.c:1231-1655
Grid point (i, j, k) sits at (i hx, j hy, k hz); arrays are indexed
[k, j, i] so that ravel() gives SPARC's x-fastest order.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .psp import Psp8, read_psp8
from .workloads import Grid, NonlocalSet

TEMP_TOL = 1e-12


def fd_d2(grid: Grid):
    """Order-`grid.order` central second-derivative weights per axis, scaled
    by 1/h^2 (src/initialization.c:2113-2147)."""
    n = grid.order // 2
    w = np.zeros(n + 1)
    for p in range(1, n + 1):
        fr = math.factorial(n) ** 2 / (math.factorial(n - p) * math.factorial(n + p))
        w[p] = 2.0 * (-1) ** (p + 1) * fr / (p * p)
    w[0] = -2.0 * sum(1.0 / (p * p) for p in range(1, n + 1))
    return w / grid.hx ** 2, w / grid.hy ** 2, w / grid.hz ** 2


def lap_box(V: np.ndarray, d2, coef: float = 1.0) -> np.ndarray:
    """coef * (FD Laplacian of V) on the interior of a box extended by FDn
    points on every side (Calc_lapV + stencil_3axis_thread_v2, cell_typ 0)."""
    dx, dy, dz = d2
    F = len(dx) - 1
    nz, ny, nx = (s - 2 * F for s in V.shape)
    c = V[F:F + nz, F:F + ny, F:F + nx]
    out = (dx[0] + dy[0] + dz[0]) * c
    for p in range(1, F + 1):
        out = out + dx[p] * (V[F:F + nz, F:F + ny, F + p:F + p + nx] + V[F:F + nz, F:F + ny, F - p:F - p + nx])
        out = out + dy[p] * (V[F:F + nz, F + p:F + p + ny, F:F + nx] + V[F:F + nz, F - p:F - p + ny, F:F + nx])
        out = out + dz[p] * (V[F + p:F + p + nz, F:F + ny, F:F + nx] + V[F - p:F - p + nz, F:F + ny, F:F + nx])
    return coef * out


def ylm(l: int, m: int, x, y, z, r):
    """Real spherical harmonics in SPARC's conventions, zero at r < 1e-10 for
    l > 0 (src/tools.c:1231-1655); l <= 3."""
    with np.errstate(invalid="ignore", divide="ignore"):
        if l == 0:
            return np.full_like(r, 0.5 * math.sqrt(1.0 / math.pi))
        if l == 1:
            v = math.sqrt(3.0 / (4.0 * math.pi)) * {-1: y, 0: z, 1: x}[m] / r
        elif l == 2:
            r2 = r * r
            v = {-2: 0.5 * math.sqrt(15.0 / math.pi) * x * y / r2,
                 -1: 0.5 * math.sqrt(15.0 / math.pi) * y * z / r2,
                 0: 0.25 * math.sqrt(5.0 / math.pi) * (-x * x - y * y + 2 * z * z) / r2,
                 1: 0.5 * math.sqrt(15.0 / math.pi) * z * x / r2,
                 2: 0.25 * math.sqrt(15.0 / math.pi) * (x * x - y * y) / r2}[m]
        elif l == 3:
            r3 = r * r * r
            v = {-3: 0.25 * math.sqrt(35.0 / (2 * math.pi)) * (3 * x * x - y * y) * y / r3,
                 -2: 0.5 * math.sqrt(105.0 / math.pi) * x * y * z / r3,
                 -1: 0.25 * math.sqrt(21.0 / (2 * math.pi)) * y * (4 * z * z - x * x - y * y) / r3,
                 0: 0.25 * math.sqrt(7.0 / math.pi) * z * (2 * z * z - 3 * x * x - 3 * y * y) / r3,
                 1: 0.25 * math.sqrt(21.0 / (2 * math.pi)) * x * (4 * z * z - x * x - y * y) / r3,
                 2: 0.25 * math.sqrt(105.0 / math.pi) * (x * x - y * y) * z / r3,
                 3: 0.25 * math.sqrt(35.0 / (2 * math.pi)) * (x * x - 3 * y * y) * x / r3}[m]
        else:
            raise ValueError("l > 3 is not supported")
    return np.where(r < 1e-10, 0.0, v)


def vloc_on(psp: Psp8, R: np.ndarray) -> np.ndarray:
    """V_loc at distances R: spline of r V_loc divided by r, V_loc(0) at the
    nucleus, -Z/r beyond the table (src/electrostatics.c:930-945)."""
    t = psp.eval("rVloc", R)
    with np.errstate(divide="ignore", invalid="ignore"):
        V = t / R
    V = np.where(R > psp.rchrg, -psp.zion / np.where(R > 0, R, 1.0), V)
    return np.where(np.abs(R) < TEMP_TOL, psp.Vloc_0, V)


def vref_on(R: np.ndarray, rcut: float, Z: float) -> np.ndarray:
    """Reference potential of a smooth charge -Z inside rcut (a degree-7
    polynomial matched to -Z/r), Calculate_Pseudopot_Ref."""
    c3 = rcut ** 3
    c6 = c3 * c3
    R2 = R * R
    inner = R2 * (R2 * R * (5.6 / c6 + R * (1.8 * R / (c6 * rcut * rcut) - 6.0 / (c6 * rcut)))
                  - 2.8 / c3) + 2.4 / rcut
    with np.errstate(divide="ignore"):
        outer = 1.0 / np.where(R > 0, R, 1.0)
    return -Z * np.where(R <= rcut, inner, outer)


def pseudocharge_cutoff(psp: Psp8, grid: Grid, tol: float) -> tuple[int, int, int]:
    """Number of grid intervals per axis of the pseudocharge box: the
    bisection of Calculate_PseudochargeCutoff on |int b + Z| < tol, b taken
    on one octant of a box around an atom at the origin."""
    d2 = fd_d2(grid)
    F = grid.order // 2
    h = np.array([grid.hx, grid.hy, grid.hz])
    rbmax = np.where(h < 1.5, h * 10 + 10, 20 * h - 9.5)
    n = np.ceil(rbmax / h - TEMP_TOL).astype(int)
    ax = [(np.arange(n[d] + 2 * F) - F) * h[d] for d in range(3)]
    Z3, Y3, X3 = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    R = np.sqrt(X3 * X3 + Y3 * Y3 + Z3 * Z3)
    b = lap_box(vloc_on(psp, R), d2, 1.0)  # [nz, ny, nx], octant i, j, k >= 0
    dV = grid.dV

    def bint(nc):
        s = b[:nc[2], :nc[1], :nc[0]].copy()
        s[0] *= 0.5
        s[nc[2] - 1] *= 0.5
        s[:, 0] *= 0.5
        s[:, nc[1] - 1] *= 0.5
        s[:, :, 0] *= 0.5
        s[:, :, nc[0] - 1] *= 0.5
        return s.sum() / (-4.0 * math.pi) * 8.0 * dV

    rc = psp.rc_max
    cur = np.array([rc, rc, rc], float)
    prev = cur.copy()
    lo = cur.copy()
    hi = rbmax.astype(float).copy()
    rtol = h.copy()
    err_cur = err_prev = 100.0
    errflag = False
    count = 0
    while count < 20:
        if err_cur < tol:
            if err_prev >= tol:
                if np.all(np.abs(cur - prev) < rtol):
                    break
                err_prev = err_cur
                tmp = cur.copy()
                hi = np.minimum(hi, cur)
                mv = np.abs(cur - prev) >= rtol
                cur = np.where(mv, 0.5 * (cur + prev), cur)
                prev = tmp
            else:
                if np.all(np.abs(cur - lo) < rtol):
                    break
                err_prev = err_cur
                prev = cur.copy()
                hi = np.minimum(hi, cur)
                mv = np.abs(cur - lo) >= rtol
                cur = np.where(mv, 0.5 * (lo + cur), cur)
        else:
            if err_prev >= tol:
                if np.all(np.abs(hi - cur) < rtol):
                    cur = hi.copy()
                    errflag = True
                    break
                err_prev = err_cur
                prev = cur.copy()
                lo = np.maximum(lo, cur)
                mv = np.abs(hi - cur) >= rtol
                cur = np.where(mv, 0.5 * (hi + cur), cur)
            else:
                if np.all(np.abs(prev - cur) < rtol):
                    cur = prev.copy()
                    err_cur = err_prev
                    break
                err_prev = err_cur
                tmp = cur.copy()
                lo = np.maximum(lo, cur)
                mv = np.abs(prev - cur) >= rtol
                cur = np.where(mv, 0.5 * (cur + prev), cur)
                prev = tmp
        nc = np.ceil(cur / h - TEMP_TOL).astype(int)
        err_cur = abs(bint(nc) + psp.zion)
        count += 1
    nc = np.ceil(cur / h - TEMP_TOL).astype(int)
    del errflag
    return int(nc[0]), int(nc[1]), int(nc[2])


def _images(x0: float, L: float, rbox: float, h: float, n: int):
    """Periodic images x0 + p L whose box [x - rbox, x + rbox] reaches the
    domain [0, (n-1) h] (the image loops of GetInfluencingAtoms_*)."""
    pmax = math.floor((rbox + L - x0) / L + TEMP_TOL)
    pmin = -math.floor((rbox + x0) / L + TEMP_TOL)
    out = []
    for p in range(pmin, pmax + 1):
        xi = x0 + L * p
        if xi < -rbox or xi >= (n - 1) * h + rbox:
            continue
        out.append(xi)
    return out


@dataclass
class AtomType:
    psp: Psp8
    positions: np.ndarray  # cartesian, Bohr, [n, 3]
    name: str = ""


@dataclass
class AtomSetup:
    grid: Grid
    types: list
    b: np.ndarray
    b_ref: np.ndarray
    Vc: np.ndarray
    Esc: float
    rho_at: np.ndarray        # guess density, integrates to nelec
    rho_core: np.ndarray      # NLCC model core density (zeros without NLCC)
    nelec: float              # -int b (the reference's PosCharge)
    n_atom: int
    cutoffs: list             # (ncx, ncy, ncz) per type
    nloc: NonlocalSet | None = None
    meta: dict = field(default_factory=dict)


def setup_atoms(grid: Grid, types: list, tol_pseudocharge: float, ref_cutoff: float = 0.5,
                nonlocal_set: bool = True) -> AtomSetup:
    """Pseudocharges, Esc, guess and core densities (Generate_PseudoChargeDensity)
    and, unless disabled, the nonlocal projectors."""
    if grid.bcx or grid.bcy or grid.bcz:
        raise NotImplementedError("setup_atoms: periodic cells only")
    d2 = fd_d2(grid)
    F = grid.order // 2
    n = (grid.nx, grid.ny, grid.nz)
    h = (grid.hx, grid.hy, grid.hz)
    L = grid.lengths
    shp = (grid.nz, grid.ny, grid.nx)
    b = np.zeros(shp); b_ref = np.zeros(shp); Vc = np.zeros(shp)
    rho_at = np.zeros(shp); rho_core = np.zeros(shp)
    esc_acc = 0.0
    inv4pi = 0.25 / math.pi
    cutoffs = []
    n_atom = 0
    for t in types:
        psp = t.psp
        nc = pseudocharge_cutoff(psp, grid, tol_pseudocharge)
        cutoffs.append(nc)
        rb = [nc[d] * h[d] for d in range(3)]
        for pos in np.asarray(t.positions, float):
            n_atom += 1
            imgs = [_images(pos[d], L[d], rb[d], h[d], n[d]) for d in range(3)]
            for z0 in imgs[2]:
                zs = max(0, math.ceil((z0 - rb[2]) / h[2])); ze = min(n[2] - 1, math.floor((z0 + rb[2]) / h[2]))
                for y0 in imgs[1]:
                    ys = max(0, math.ceil((y0 - rb[1]) / h[1])); ye = min(n[1] - 1, math.floor((y0 + rb[1]) / h[1]))
                    for x0 in imgs[0]:
                        xs = max(0, math.ceil((x0 - rb[0]) / h[0])); xe = min(n[0] - 1, math.floor((x0 + rb[0]) / h[0]))
                        if xs > xe or ys > ye or zs > ze:
                            continue
                        ax = np.arange(xs - F, xe + F + 1) * h[0] - x0
                        ay = np.arange(ys - F, ye + F + 1) * h[1] - y0
                        az = np.arange(zs - F, ze + F + 1) * h[2] - z0
                        Zg, Yg, Xg = np.meshgrid(az, ay, ax, indexing="ij")
                        R = np.sqrt(Xg * Xg + Yg * Yg + Zg * Zg)
                        VJ = vloc_on(psp, R)
                        VJr = vref_on(R, ref_cutoff, psp.zion)
                        inner = (slice(F, F + ze - zs + 1), slice(F, F + ye - ys + 1),
                                 slice(F, F + xe - xs + 1))
                        Ri = R[inner]
                        rJ = np.where(Ri > psp.rchrg, 0.0, psp.eval("rhoIsoAtom", Ri))
                        rcJ = np.where(Ri > psp.rchrg, 0.0, psp.eval("rho_c_table", Ri))
                        bJ = lap_box(VJ, d2, -inv4pi)
                        bJr = lap_box(VJr, d2, -inv4pi)
                        dst = (slice(zs, ze + 1), slice(ys, ye + 1), slice(xs, xe + 1))
                        b[dst] += bJ
                        b_ref[dst] += bJr
                        Vc[dst] += VJr[inner] - VJ[inner]
                        rho_at[dst] += rJ
                        rho_core[dst] += rcJ
                        esc_acc -= float((bJr * VJr[inner]).sum())
    dV = grid.dV
    esc = 0.5 * dV * (esc_acc + float(((b + b_ref) * Vc).sum()))
    pos_charge = -b.sum() * dV
    neg_charge = -rho_at.sum() * dV
    rho_at *= (0.0 - pos_charge) / neg_charge
    out = AtomSetup(grid, types, b.ravel(), b_ref.ravel(), Vc.ravel(), esc, rho_at.ravel(),
                    rho_core.ravel(), float(pos_charge), n_atom, cutoffs)
    if nonlocal_set:
        out.nloc = nonlocal_projectors(grid, types)
    return out


def nonlocal_projectors(grid: Grid, types: list) -> NonlocalSet:
    """Chi = UdV_spline(r) Y_lm on the grid points within rc of every periodic
    image; all images of an atom share its alpha rows (IP_displ).  Projector
    order per atom: l ascending (skipping lloc), radial index, m = -l..l."""
    n = (grid.nx, grid.ny, grid.nz)
    h = (grid.hx, grid.hy, grid.hz)
    L = grid.lengths
    ndc, nproj, pos_off, chi_off, alpha_off = [], [], [], [], []
    pos_all, chi_all, gam_all = [], [], []
    po = co = ao = 0
    for t in types:
        psp = t.psp
        rc = psp.rc_max
        npj = psp.nproj_total
        if npj == 0:
            continue
        lpos = np.concatenate([[0], np.cumsum(psp.ppl)])
        gam = []
        for l in range(psp.lmax + 1):
            if l == psp.lloc:
                continue
            for p in range(psp.ppl[l]):
                gam += [psp.Gamma[lpos[l] + p]] * (2 * l + 1)
        for pos in np.asarray(t.positions, float):
            atom_ao = ao
            nimg = 0
            imgs = [_images(pos[d], L[d], rc, h[d], n[d]) for d in range(3)]
            for z0 in imgs[2]:
                zs = max(0, math.ceil((z0 - rc) / h[2])); ze = min(n[2] - 1, math.floor((z0 + rc) / h[2]))
                for y0 in imgs[1]:
                    ys = max(0, math.ceil((y0 - rc) / h[1])); ye = min(n[1] - 1, math.floor((y0 + rc) / h[1]))
                    for x0 in imgs[0]:
                        xs = max(0, math.ceil((x0 - rc) / h[0])); xe = min(n[0] - 1, math.floor((x0 + rc) / h[0]))
                        if xs > xe or ys > ye or zs > ze:
                            continue
                        K, J, I = np.meshgrid(np.arange(zs, ze + 1), np.arange(ys, ye + 1),
                                              np.arange(xs, xe + 1), indexing="ij")
                        dx = (I * h[0] - x0).ravel(); dy = (J * h[1] - y0).ravel()
                        dz = (K * h[2] - z0).ravel()
                        r2 = dx * dx + dy * dy + dz * dz
                        sel = r2 <= rc * rc
                        if not sel.any():
                            continue
                        dx, dy, dz = dx[sel], dy[sel], dz[sel]
                        lin = (K.ravel()[sel] * n[1] + J.ravel()[sel]) * n[0] + I.ravel()[sel]
                        r = np.sqrt(dx * dx + dy * dy + dz * dz)
                        cols = []
                        for l in range(psp.lmax + 1):
                            if l == psp.lloc:
                                continue
                            for p in range(psp.ppl[l]):
                                u = psp.eval("UdV", r, lpos[l] + p)
                                for m in range(-l, l + 1):
                                    cols.append(ylm(l, m, dx, dy, dz, r) * u)
                        C = np.stack(cols, axis=1)
                        ndc.append(lin.size); nproj.append(npj)
                        pos_off.append(po); chi_off.append(co); alpha_off.append(atom_ao)
                        pos_all.append(lin.astype(np.int32))
                        chi_all.append(C.T.ravel())
                        po += lin.size; co += C.size
                        nimg += 1
            if nimg:
                ao += npj
                gam_all.extend(gam)
    return NonlocalSet(
        ndc=np.asarray(ndc, np.int32), nproj=np.asarray(nproj, np.int32),
        pos_off=np.asarray(pos_off, np.int64),
        grid_pos=np.concatenate(pos_all) if pos_all else np.zeros(0, np.int32),
        chi_off=np.asarray(chi_off, np.int64),
        chi=np.concatenate(chi_all) if chi_all else np.zeros(0),
        alpha_off=np.asarray(alpha_off, np.int32), n_alpha=ao,
        gamma=np.asarray(gam_all, np.float64), dV=grid.dV)


def load_types(spec) -> list:
    """[(psp8 file, positions), ...] -> [AtomType, ...]."""
    return [AtomType(read_psp8(f), np.asarray(p, float), name=str(f)) for f, p in spec]
