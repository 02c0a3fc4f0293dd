"""CPU restatement of the atomic force terms, FP64 numpy:

  * nonlocal (Calculate_nonlocal_forces_linear, src/forces.c:750-847;
    Compute_Integral_psi_Chi :849-888; Compute_Integral_Chi_Dpsi :890-924;
    Compute_force_nloc_by_integrals :926-993): the first derivative of every
    orbital with the order-2n central stencil (Gradient_vectors_dir, periodic
    wrap-around), the gather of each image atom's sphere points and the
    products with its Chi, images summed into the atom's rows;
  * local (Calculate_local_forces_linear, :157-457) and NLCC
    (Calculate_forces_xc_linear, :473-707) over each image atom's
    pseudocharge box, with the set-up's splines and FD Laplacian:
        dV sum [ -b_J grad(phi) + 1/2 ( grad(Vc_J) (b + b_ref)
                                       - grad(Vc) (b_J + b_J^ref) ) ]
        dV sum Vxc grad(rho_c,J)

TEST INFRASTRUCTURE ONLY: the checker of sgpu_nonlocal_integrals and
sgpu_local_forces.
"""
from __future__ import annotations

import math

import numpy as np

from sparc_amd.atoms import AtomSetup, fd_d2, lap_box, vloc_on, vref_on
from sparc_amd.forces import _atom_boxes


def d1_weights(order: int, h: float) -> np.ndarray:
    """w[p] = (-1)^(p+1) n!^2 / ((n-p)! (n+p)!) / (p h), w[0] = 0
    (src/initialization.c:2113-2132)."""
    n = order // 2
    w = np.zeros(n + 1)
    for p in range(1, n + 1):
        w[p] = (-1) ** (p + 1) * math.factorial(n) ** 2 / (
            math.factorial(n - p) * math.factorial(n + p)) / p / h
    return w


def gradient(grid, X: np.ndarray, axis: int) -> np.ndarray:
    """d/dx (axis 0), d/dy (1), d/dz (2) of every column of X (Nd x Ns,
    x fastest), periodic."""
    h = (grid.hx, grid.hy, grid.hz)[axis]
    w = d1_weights(grid.order, h)
    ns = X.shape[1]
    V = X.T.reshape(ns, grid.nz, grid.ny, grid.nx)
    ax = 3 - axis
    out = np.zeros_like(V)
    for p in range(1, len(w)):
        out += w[p] * (np.roll(V, -p, axis=ax) - np.roll(V, p, axis=ax))
    return out.reshape(ns, -1).T


def integrals(grid, nloc, X: np.ndarray) -> np.ndarray:
    """[4, n_alpha, Ns]: dV Chi' psi and Chi' d_x psi, Chi' d_y psi,
    Chi' d_z psi, images summed into their atom's rows."""
    X = np.asarray(X, np.float64)
    ns = X.shape[1]
    out = np.zeros((4, nloc.n_alpha, ns))
    D = [X] + [gradient(grid, X, a) for a in range(3)]
    for a in range(nloc.n_influence):
        n, p = int(nloc.ndc[a]), int(nloc.nproj[a])
        pos = nloc.grid_pos[nloc.pos_off[a]:nloc.pos_off[a] + n]
        C = nloc.chi[nloc.chi_off[a]:nloc.chi_off[a] + n * p].reshape(p, n)
        r0 = int(nloc.alpha_off[a])
        for k in range(4):
            s = nloc.dV if k == 0 else 1.0
            out[k, r0:r0 + p] += s * (C @ D[k][pos])
    return out


def nonlocal_forces(grid, nloc, atom_rows, n_atom: int, X: np.ndarray, occ: np.ndarray,
                    occfac: float) -> np.ndarray:
    """F^nl [n_atom, 3] = -2 occfac sum_n g_n sum_p Gamma_p alpha_pn beta_pn;
    atom_rows: (atom index, first alpha row, projector count) per atom."""
    I = integrals(grid, nloc, X)
    f = np.zeros((n_atom, 3))
    g = np.asarray(occ, float)
    for ia, r0, npj in atom_rows:
        gam = nloc.gamma[r0:r0 + npj, None]
        for d in range(3):
            f[ia, d] = -2.0 * occfac * float(((gam * I[0, r0:r0 + npj] * I[1 + d, r0:r0 + npj]).sum(0) * g).sum())
    return f


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
