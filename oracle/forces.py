"""CPU restatement of the nonlocal force integrals and their combination
(Calculate_nonlocal_forces_linear, src/forces.c:750-847;
Compute_Integral_psi_Chi :849-888; Compute_Integral_Chi_Dpsi :890-924;
Compute_force_nloc_by_integrals :926-993), FP64 numpy: the first derivative
of every orbital with the order-2n central stencil (Gradient_vectors_dir,
periodic wrap-around), the gather of each image atom's sphere points and the
products with its Chi, images summed into the atom's rows.

TEST INFRASTRUCTURE ONLY: the checker of sgpu_nonlocal_integrals.
"""
from __future__ import annotations

import math

import numpy as np


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
