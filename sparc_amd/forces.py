"""Atomic forces after a converged SCF (SPARC Calculate_EGS_Forces,
src/forces.c:49-142), for the cells the device SCF runs: orthogonal,
periodic, spin-unpolarized, Gamma point, LDA.

    F_J = F_J^loc + F_J^nl + F_J^xc, then the mean over atoms is removed
    (Symmetrize_forces, src/forces.c:1399-1427).

  * F^loc (Calculate_local_forces_linear, src/forces.c:157-457), per image
    atom J over its pseudocharge box (the boxes of the set-up):
        dV sum [ -b_J grad(phi) + 1/2 ( grad(Vc_J) (b + b_ref)
                                       - grad(Vc) (b_J + b_J^ref) ) ]
    with b_J = -Lap(V_J) / 4 pi, b_J^ref = -Lap(V_J^ref) / 4 pi and
    Vc_J = V_J^ref - V_J; host code over the atom boxes, like the set-up.
  * F^xc (Calculate_forces_xc_linear, src/forces.c:473-707), NLCC only:
        dV sum Vxc grad(rho_c,J)    over the same boxes; host code.
  * F^nl (Calculate_nonlocal_forces_linear, src/forces.c:750-847 with
    Compute_force_nloc_by_integrals, :926-993):
        -2 occfac sum_n g_n sum_p Gamma_p <Chi_p, psi_n> dV <Chi_p, d psi_n>
    summed over each atom's projectors p; the integrals over all orbitals
    are formed on the device (sgpu_nonlocal_integrals: the order-12
    first-derivative stencil of every orbital and the projections, images
    summed), the O(n_alpha Ns) combination here.

Gradients use the reference's first-derivative weights
w_p = (-1)^(p+1) fract(n, p) / (p h) (src/initialization.c:2113-2132),
periodic wrap-around for the global fields.
"""
from __future__ import annotations

import math

import numpy as np

from .atoms import AtomSetup, _images, lap_box, vloc_on, vref_on, fd_d2, pseudocharge_cutoff


def fd_d1(order: int, h: float) -> np.ndarray:
    """First-derivative weights w[1..n] (w[0] = 0) scaled by 1/h."""
    n = order // 2
    w = np.zeros(n + 1)
    for p in range(1, n + 1):
        fr = math.factorial(n) ** 2 / (math.factorial(n - p) * math.factorial(n + p))
        w[p] = (-1) ** (p + 1) * fr / p
    return w / h


def grad_periodic(f: np.ndarray, w: np.ndarray, axis: int) -> np.ndarray:
    """Central FD derivative of a periodic [nz, ny, nx] field along axis
    (0 z, 1 y, 2 x)."""
    out = np.zeros_like(f)
    for p in range(1, len(w)):
        out += w[p] * (np.roll(f, -p, axis=axis) - np.roll(f, p, axis=axis))
    return out


def grad_box(V: np.ndarray, w: tuple, F: int) -> tuple:
    """(d/dx, d/dy, d/dz) of V on the interior of a box extended by F points
    on every side ([z, y, x] arrays)."""
    nz, ny, nx = (s - 2 * F for s in V.shape)
    wx, wy, wz = w
    gx = np.zeros((nz, ny, nx)); gy = np.zeros_like(gx); gz = np.zeros_like(gx)
    cz, cy, cx = slice(F, F + nz), slice(F, F + ny), slice(F, F + nx)
    for p in range(1, F + 1):
        gx += wx[p] * (V[cz, cy, F + p:F + p + nx] - V[cz, cy, F - p:F - p + nx])
        gy += wy[p] * (V[cz, F + p:F + p + ny, cx] - V[cz, F - p:F - p + ny, cx])
        gz += wz[p] * (V[F + p:F + p + nz, cy, cx] - V[F - p:F - p + nz, cy, cx])
    return gx, gy, gz


def _atom_boxes(setup: AtomSetup, tol_pseudocharge: float):
    """(atom index, type, image centre, inner slices of the global grid, the
    F-extended box coordinates) over every image atom's pseudocharge box, in
    the order of the set-up (GetInfluencingAtoms_local)."""
    grid = setup.grid
    F = grid.order // 2
    n = (grid.nx, grid.ny, grid.nz)
    h = (grid.hx, grid.hy, grid.hz)
    L = grid.lengths
    ia = 0
    for ti, t in enumerate(setup.types):
        nc = setup.cutoffs[ti] if setup.cutoffs else pseudocharge_cutoff(t.psp, grid, tol_pseudocharge)
        rb = [nc[d] * h[d] for d in range(3)]
        for pos in np.asarray(t.positions, float):
            imgs = [_images(pos[d], L[d], rb[d], h[d], n[d]) for d in range(3)]
            for z0 in imgs[2]:
                zs = max(0, math.ceil((z0 - rb[2]) / h[2])); ze = min(n[2] - 1, math.floor((z0 + rb[2]) / h[2]))
                for y0 in imgs[1]:
                    ys = max(0, math.ceil((y0 - rb[1]) / h[1])); ye = min(n[1] - 1, math.floor((y0 + rb[1]) / h[1]))
                    for x0 in imgs[0]:
                        xs = max(0, math.ceil((x0 - rb[0]) / h[0])); xe = min(n[0] - 1, math.floor((x0 + rb[0]) / h[0]))
                        if xs > xe or ys > ye or zs > ze:
                            continue
                        dst = (slice(zs, ze + 1), slice(ys, ye + 1), slice(xs, xe + 1))
                        yield ia, t, (x0, y0, z0), dst, (xs, ys, zs, xe, ye, ze)
            ia += 1


def _box_R(grid, c0, lim, ext):
    x0, y0, z0 = c0
    xs, ys, zs, xe, ye, ze = lim
    ax = np.arange(xs - ext, xe + ext + 1) * grid.hx - x0
    ay = np.arange(ys - ext, ye + ext + 1) * grid.hy - y0
    az = np.arange(zs - ext, ze + ext + 1) * grid.hz - z0
    Zg, Yg, Xg = np.meshgrid(az, ay, ax, indexing="ij")
    return np.sqrt(Xg * Xg + Yg * Yg + Zg * Zg)


def local_forces(setup: AtomSetup, phi: np.ndarray, ref_cutoff: float = 0.5,
                 tol_pseudocharge: float = 1e-8) -> np.ndarray:
    """F^loc [n_atom, 3] from the electrostatic potential phi (Nd)."""
    grid = setup.grid
    F = grid.order // 2
    shp = (grid.nz, grid.ny, grid.nx)
    w = (fd_d1(grid.order, grid.hx), fd_d1(grid.order, grid.hy), fd_d1(grid.order, grid.hz))
    d2 = fd_d2(grid)
    inv4pi = 0.25 / math.pi
    ph = np.asarray(phi, float).reshape(shp)
    Vc = np.asarray(setup.Vc, float).reshape(shp)
    bsum = (np.asarray(setup.b, float) + np.asarray(setup.b_ref, float)).reshape(shp)
    dphi = [grad_periodic(ph, w[2 - a], a) for a in (2, 1, 0)]     # x, y, z
    dVc = [grad_periodic(Vc, w[2 - a], a) for a in (2, 1, 0)]
    f = np.zeros((setup.n_atom, 3))
    for ia, t, c0, dst, lim in _atom_boxes(setup, tol_pseudocharge):
        R = _box_R(grid, c0, lim, F)
        VJ = vloc_on(t.psp, R)
        VJr = vref_on(R, ref_cutoff, t.psp.zion)
        bJ = lap_box(VJ, d2, -inv4pi)
        bJr = lap_box(VJr, d2, -inv4pi)
        gVcJ = grad_box(VJr - VJ, w, F)
        bb = bsum[dst]
        for d in range(3):
            fl = -float((bJ * dphi[d][dst]).sum())
            fc = float((gVcJ[d] * bb).sum() - (dVc[d][dst] * (bJ + bJr)).sum())
            f[ia, d] += (fl + 0.5 * fc) * grid.dV
    return f


def xc_forces(setup: AtomSetup, vxc: np.ndarray, tol_pseudocharge: float = 1e-8) -> np.ndarray:
    """F^xc [n_atom, 3] of the NLCC core densities (zero without NLCC)."""
    grid = setup.grid
    F = grid.order // 2
    shp = (grid.nz, grid.ny, grid.nx)
    w = (fd_d1(grid.order, grid.hx), fd_d1(grid.order, grid.hy), fd_d1(grid.order, grid.hz))
    v = np.asarray(vxc, float).reshape(shp)
    f = np.zeros((setup.n_atom, 3))
    for ia, t, c0, dst, lim in _atom_boxes(setup, tol_pseudocharge):
        psp = t.psp
        if not np.any(psp.rho_c_table):
            continue
        R = _box_R(grid, c0, lim, F)
        rc = np.where(R > psp.rchrg, 0.0, psp.eval("rho_c_table", R))
        g = grad_box(rc, w, F)
        vv = v[dst]
        for d in range(3):
            f[ia, d] += float((vv * g[d]).sum()) * grid.dV
    return f


def atom_rows(setup: AtomSetup):
    """(atom index, first alpha row, projector count) of every atom with
    projectors, in the order of the projector set (IP_displ)."""
    out = []
    ia = ao = 0
    for t in setup.types:
        npj = t.psp.nproj_total
        for _ in np.asarray(t.positions, float):
            if npj:
                out.append((ia, ao, npj))
                ao += npj
            ia += 1
    return out


def nonlocal_forces_from_integrals(setup: AtomSetup, ints: np.ndarray, occ: np.ndarray,
                                   occfac: float) -> np.ndarray:
    """Compute_force_nloc_by_integrals: ints [4, n_alpha, Ns] holds
    dV <Chi, psi> and <Chi, d_x psi>, <Chi, d_y psi>, <Chi, d_z psi>."""
    gam = np.asarray(setup.nloc.gamma, float)
    g = np.asarray(occ, float)
    f = np.zeros((setup.n_atom, 3))
    for ia, r0, npj in atom_rows(setup):
        a = ints[0, r0:r0 + npj]                       # [npj, Ns]
        for d in range(3):
            s = (gam[r0:r0 + npj, None] * a * ints[1 + d, r0:r0 + npj]).sum(axis=0)
            f[ia, d] = -2.0 * occfac * float((s * g).sum())
    return f


def symmetrize(f: np.ndarray) -> np.ndarray:
    return f - f.mean(axis=0, keepdims=True)


def nonlocal_forces(ctx, setup: AtomSetup, psi, occ: np.ndarray, occfac: float) -> np.ndarray:
    """F^nl [n_atom, 3] from the device orbitals psi (a Block of ctx, whose
    projectors are setup.nloc)."""
    if setup.nloc is None or setup.nloc.n_alpha == 0:
        return np.zeros((setup.n_atom, 3))
    return nonlocal_forces_from_integrals(setup, ctx.nonlocal_integrals(psi), occ, occfac)


def scf_forces(ctx, setup: AtomSetup, params: dict, result: dict, ref_cutoff: float = 0.5,
               tol_pseudocharge: float = 1e-8) -> dict:
    """Forces after ctx.scf_run(..., keep_psi=True): the three parts and
    their symmetrized sum (Ha/Bohr, [n_atom, 3] each)."""
    from .sparc_io import occupations
    occ = occupations(result["lambda_"], result["Efermi"], params["beta"], params["smear_type"])
    floc = local_forces(setup, result["phi"], ref_cutoff, tol_pseudocharge)
    fnl = nonlocal_forces(ctx, setup, result["psi"], occ, params["occfac"])
    fxc = xc_forces(setup, result["vxc"], tol_pseudocharge)
    return {"local": floc, "nonlocal": fnl, "xc": fxc, "total": symmetrize(floc + fnl + fxc)}
