"""Shared builders for the parity tests (small seeded cases)."""
from __future__ import annotations

import numpy as np

from sparc_amd.workloads import Grid, NonlocalSet, gaussian_projectors, gaussian_veff


def small_case(n=(14, 13, 12), h=(0.45, 0.5, 0.55), bc=(0, 0, 0), n_sites=3, rc=2.2, seed=0,
               order=12, lmax=1):
    g = Grid(*n, *h, *bc, order=order)
    rng = np.random.default_rng(seed)
    veff = gaussian_veff(g, 4, rng)
    L = np.asarray(g.lengths)
    sites = rng.uniform(0.2, 0.8, size=(n_sites, 3)) * L if any(bc) else \
        rng.uniform(0, 1, size=(n_sites, 3)) * L
    rc = min(rc, 0.45 * float(min(L)))
    nl = gaussian_projectors(g, sites, rc, sigmas=(1.0, 0.7), gammas=(1.3, -0.6), lmax=lmax)
    return g, veff, nl, rng


def split_images(nl: NonlocalSet, entry: int = 0) -> NonlocalSet:
    """Split one influence entry into two entries that share alpha rows, the
    way SPARC lists periodic images of one atom as separate influence atoms."""
    a = entry
    n, p = int(nl.ndc[a]), int(nl.nproj[a])
    pos = nl.grid_pos[nl.pos_off[a]:nl.pos_off[a] + n]
    C = nl.chi[nl.chi_off[a]:nl.chi_off[a] + n * p].reshape(p, n)
    h = n // 2
    parts = [(pos[:h], C[:, :h]), (pos[h:], C[:, h:])]
    ndc, nproj, pos_off, chi_off, alpha_off, gp, ch = [], [], [], [], [], [], []
    po = co = 0
    for b in range(nl.n_influence):
        if b == a:
            items = parts
        else:
            nb, pb = int(nl.ndc[b]), int(nl.nproj[b])
            items = [(nl.grid_pos[nl.pos_off[b]:nl.pos_off[b] + nb],
                      nl.chi[nl.chi_off[b]:nl.chi_off[b] + nb * pb].reshape(pb, nb))]
        for ps, cs in items:
            ndc.append(ps.size); nproj.append(cs.shape[0]); pos_off.append(po); chi_off.append(co)
            alpha_off.append(int(nl.alpha_off[b]))
            gp.append(ps); ch.append(cs.ravel())
            po += ps.size; co += cs.size
    return NonlocalSet(np.asarray(ndc, np.int32), np.asarray(nproj, np.int32),
                       np.asarray(pos_off, np.int64), np.concatenate(gp).astype(np.int32),
                       np.asarray(chi_off, np.int64), np.concatenate(ch), np.asarray(alpha_off, np.int32),
                       nl.n_alpha, nl.gamma.copy(), nl.dV)


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
