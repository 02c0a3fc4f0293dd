"""Atomic forces after a converged SCF (SPARC Calculate_EGS_Forces,
src/forces.c:49-142), for the cells the device SCF runs: orthogonal,
periodic, spin-unpolarized, Gamma point, LDA.

    F_J = F_J^loc + F_J^nl + F_J^xc, then the mean over atoms is removed
    (Symmetrize_forces, src/forces.c:1399-1427).

  * F^loc (Calculate_local_forces_linear, src/forces.c:157-457) and F^xc
    (Calculate_forces_xc_linear, :473-707), per image atom J over its
    pseudocharge box:
        dV sum [ -b_J grad(phi) + 1/2 ( grad(Vc_J) (b + b_ref)
                                       - grad(Vc) (b_J + b_J^ref) ) ]
        dV sum Vxc grad(rho_c,J)     (NLCC only)
    with b_J = -Lap(V_J) / 4 pi, b_J^ref = -Lap(V_J^ref) / 4 pi and
    Vc_J = V_J^ref - V_J: on the device (sgpu_local_forces); the host lists
    the image boxes (the set-up's GetInfluencingAtoms_local boxes) and the
    radial tables.
  * F^nl (Calculate_nonlocal_forces_linear, src/forces.c:750-847 with
    Compute_force_nloc_by_integrals, :926-993):
        -2 occfac sum_n g_n sum_p Gamma_p <Chi_p, psi_n> dV <Chi_p, d psi_n>
    summed over each atom's projectors p; the integrals over all orbitals
    are formed on the device (sgpu_nonlocal_integrals: the order-12
    first-derivative stencil of every orbital and the projections, images
    summed), the O(n_alpha Ns) combination here.
"""
from __future__ import annotations

import math

import numpy as np

from .atoms import AtomSetup, _images, pseudocharge_cutoff


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


def image_boxes(setup: AtomSetup, tol_pseudocharge: float = 1e-8):
    """(type index, atom index, box xs ys zs xe ye ze, centre) per image atom."""
    tidx = {id(t): i for i, t in enumerate(setup.types)}
    types, atoms, boxes, centres = [], [], [], []
    for ia, t, c0, dst, lim in _atom_boxes(setup, tol_pseudocharge):
        types.append(tidx[id(t)])
        atoms.append(ia)
        boxes.append(lim)
        centres.append(c0)
    return (np.asarray(types, np.int32), np.asarray(atoms, np.int32),
            np.asarray(boxes, np.int32).reshape(-1, 6), np.asarray(centres, float).reshape(-1, 3))


def local_forces(ctx, setup: AtomSetup, phi: np.ndarray, vxc: np.ndarray, ref_cutoff: float = 0.5,
                 tol_pseudocharge: float = 1e-8):
    """(F^loc, F^xc) [n_atom, 3] each, on the device (sgpu_local_forces)."""
    return ctx.local_forces(setup, image_boxes(setup, tol_pseudocharge), phi, vxc, ref_cutoff)


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
    floc, fxc = local_forces(ctx, setup, result["phi"], result["vxc"], ref_cutoff, tol_pseudocharge)
    fnl = nonlocal_forces(ctx, setup, result["psi"], occ, params["occfac"])
    return {"local": floc, "nonlocal": fnl, "xc": fxc, "total": symmetrize(floc + fnl + fxc)}
