"""FP64 numpy restatement of the symmetric tridiagonal eigensolver by divide
and conquer (Cuppen's method with LAPACK's rules: dstedc / dlaed0-dlaed4),
which the reference reaches through LAPACKE_dsygvd (src/eigenSolver.c:1318:
dsygvd -> dsyevd -> dstedc) and which sparc_amd/csrc/tdc.hip runs on the
device.

TEST INFRASTRUCTURE ONLY: used by tests/ to pin the algorithm the device
kernels follow; pure-Python loops, for small matrices.

    tdc(d, e) -> (lam ascending, Z) with T = Z diag(lam) Z', T the
    tridiagonal matrix of diagonal d and off-diagonal e.

Steps (the same as the device's): 1 x 1 leaves with d_i - |e_i-1| - |e_i|;
merges of [a, m) and [m, b) with beta = e[m-1]: z = (last row of the left
eigenvectors, sign(beta) first row of the right ones), rho = |beta| |z|^2,
z normalised; the two sorted runs merged by rank; deflation in sorted order
(dlaed2): rho |z_j| <= tol deflates pole j, and of two consecutive kept poles
whose Givens rotation (c, s zeroing the first z) has |t c s| <= tol
(t the gap) the first deflates and the second, rotated, is tested against
the next; tol = 8 eps max(|d|max, |z|max); the secular equation
1/rho + sum z_j^2 / (d_j - lambda) = 0 solved per root from the nearer pole
by Li's middle-way iteration (LAPACK dlaed4) inside a bracket; the vectors
from Gu and Eisenstat's recomputed z (dlaed3); roots and deflated values
merged in ascending order.

deflate_scan is dlaed2's sequential scan; deflate_chains evaluates the same
decisions the way the device does (every consecutive kept pair tested on
the original values, one chain walk per pair that passes, a chain starting
inside an earlier one dropped) and must agree with it exactly.
"""
from __future__ import annotations

import numpy as np

EPS = np.finfo(float).eps


def _pair_test(dp, zp, dn, zn, tol):
    tau = np.hypot(zn, zp)
    c, s = zn / tau, -zp / tau
    return abs((dn - dp) * c * s) <= tol, c, s, tau


def deflate_scan(ds, zs, rho, tol):
    """dlaed2's scan in sorted order.  Returns (deflated indices, kept
    indices, rotations [(first, second, c, s)], poles, z values, deflated
    values by index)."""
    ds, zs = ds.copy(), zs.copy()
    nl = len(ds)
    if rho * np.max(np.abs(zs)) <= tol:
        return list(range(nl)), [], [], ds, zs, {j: ds[j] for j in range(nl)}
    df, nd, rots, dval = [], [], [], {}
    pj = None
    for j in range(nl):
        if rho * abs(zs[j]) <= tol:
            df.append(j)
            dval[j] = ds[j]
            continue
        if pj is None:
            pj = j
            continue
        ok, c, s, tau = _pair_test(ds[pj], zs[pj], ds[j], zs[j], tol)
        if ok:
            zs[j], zs[pj] = tau, 0.0
            rots.append((pj, j, c, s))
            t2 = ds[pj] * c * c + ds[j] * s * s
            ds[j] = ds[pj] * s * s + ds[j] * c * c
            ds[pj] = t2
            df.append(pj)
            dval[pj] = t2
        else:
            nd.append(pj)
        pj = j
    if pj is not None:
        nd.append(pj)
    return df, nd, rots, ds, zs, dval


def deflate_chains(ds, zs, rho, tol):
    """The same decisions as deflate_scan, evaluated as the device does."""
    ds, zs = ds.copy(), zs.copy()
    nl = len(ds)
    small = (rho * np.abs(zs) <= tol) | (rho * np.max(np.abs(zs)) <= tol)
    L = [j for j in range(nl) if not small[j]]
    M = len(L)
    cand = [k for k in range(M - 1)
            if _pair_test(ds[L[k]], zs[L[k]], ds[L[k + 1]], zs[L[k + 1]], tol)[0]]
    rotated, newv = {}, {}

    def walk(k, write):
        dp, zp = ds[L[k]], zs[L[k]]
        while k < M - 1:
            j = L[k + 1]
            ok, c, s, tau = _pair_test(dp, zp, ds[j], zs[j], tol)
            if not ok:
                return k
            t2 = dp * c * c + ds[j] * s * s
            dn = dp * s * s + ds[j] * c * c
            if write:
                rotated[k] = (c, s, t2)
                newv[j] = (dn, tau)
            dp, zp = dn, tau
            k += 1
        return k

    ends = [walk(c, False) for c in cand]
    last = -1
    valid = []
    for c, e in zip(cand, ends):
        if c > last:
            valid.append(c)
            last = e
    for c in valid:
        walk(c, True)
    for j, (dn, tau) in newv.items():
        ds[j], zs[j] = dn, tau
    nd = [L[k] for k in range(M) if not (k < M - 1 and k in rotated)]
    df = [j for j in range(nl) if small[j]] + [L[k] for k in sorted(rotated)]
    dval = {j: ds[j] for j in range(nl) if small[j]}
    for k, (c, s, t2) in rotated.items():
        dval[L[k]] = t2
    rots = [(L[k], L[k + 1], rotated[k][0], rotated[k][1]) for k in sorted(rotated)]
    return df, nd, rots, ds, zs, dval


def secular_root(dd, zz, rho, i, max_iter=200):
    """Root i of 1/rho + sum zz_j^2 / (dd_j - lambda) = 0 (dd ascending):
    (origin pole o, tau) with lambda = o + tau, by the middle-way iteration."""
    K = len(dd)
    w = zz * zz
    left = np.arange(K) <= i

    def f(delta, t):
        return 1.0 / rho + np.sum(w / (delta - t))

    if i < K - 1:
        mid = 0.5 * (dd[i + 1] - dd[i])
        if f(dd - dd[i], mid) > 0:
            o, lo, hi, t = dd[i], 0.0, mid, mid
        else:
            o, lo, hi, t = dd[i + 1], -mid, 0.0, -mid
    else:
        o, lo, hi = dd[K - 1], 0.0, rho * np.sum(w)
        t = hi
    delta = dd - o
    for _ in range(max_iter):
        r = 1.0 / (delta - t)
        psi, phi = np.sum((w * r)[left]), np.sum((w * r)[~left])
        dpsi, dphi = np.sum((w * r * r)[left]), np.sum((w * r * r)[~left])
        g = 1.0 / rho + psi + phi
        if abs(g) <= EPS * (8.0 * (phi - psi + 1.0 / rho) + abs(t) * (dpsi + dphi)):
            break
        if g > 0:
            hi = t
        else:
            lo = t
        if not hi > lo or hi - lo <= 2 * EPS * max(abs(lo), abs(hi)):
            break
        a1 = delta[i] - t
        B = dpsi * a1 * a1
        tn = None
        if i < K - 1:
            a2 = delta[i + 1] - t
            D = dphi * a2 * a2
            c0 = 1.0 / rho + (psi - B / a1) + (phi - D / a2)
            bq = c0 * (a1 + a2) + B + D        # c0 eta^2 - bq eta + a1 a2 g = 0
            r0 = a1 * a2 * g
            disc = bq * bq - 4.0 * c0 * r0
            if disc >= 0:
                q = 0.5 * (bq + np.copysign(np.sqrt(disc), bq))
                for eta in ([r0 / q] if q != 0 else []) + ([q / c0] if c0 != 0 else []):
                    if lo < t + eta < hi:
                        tn = t + eta
                        break
        else:
            c0 = 1.0 / rho + (psi - B / a1)
            if c0 > 0 and lo < t + a1 + B / c0 < hi:
                tn = t + a1 + B / c0
        if tn is None or not lo < tn < hi:
            tn = 0.5 * (lo + hi)
        if tn == t:
            break
        t = tn
    return o, t


def tdc(d, e, deflate=deflate_chains, root=secular_root):
    """Eigenvalues (ascending) and eigenvectors of the tridiagonal matrix."""
    n = len(d)
    lam = np.array(d, dtype=float)
    for i in range(n - 1):
        lam[i] -= abs(e[i])
        lam[i + 1] -= abs(e[i])
    Z = np.eye(n)
    s = 1
    while s < n:
        lamn, Zn = lam.copy(), Z.copy()
        for a in range(0, n, 2 * s):
            m = a + s
            if m >= n:
                continue
            b = min(a + 2 * s, n)
            beta = e[m - 1]
            sg = 1.0 if beta >= 0 else -1.0
            z = np.array([Z[m - 1, j] for j in range(a, m)] + [sg * Z[m, j] for j in range(m, b)])
            nz = np.sqrt(np.sum(z * z))
            z = z / nz
            rho = abs(beta) * nz * nz
            dl = lam[a:b]
            lo_run, hi_run = dl[:m - a], dl[m - a:]
            pos = np.zeros(b - a, int)
            for i in range(m - a):
                pos[i] = i + np.searchsorted(hi_run, lo_run[i], side="left")
            for j in range(b - m):
                pos[m - a + j] = j + np.searchsorted(lo_run, hi_run[j], side="right")
            perm = np.empty(b - a, int)
            perm[pos] = np.arange(b - a)
            ds, zs = dl[perm].copy(), z[perm].copy()
            cols = [a + p for p in perm]
            Zb = Z[a:b, :].copy()
            tol = 8 * EPS * max(np.max(np.abs(ds)), np.max(np.abs(zs)))
            df, nd, rots, ds, zs, dval = deflate(ds, zs, rho, tol)
            for pj, j, c, sn in rots:
                x, y = Zb[:, cols[pj]].copy(), Zb[:, cols[j]].copy()
                Zb[:, cols[pj]] = c * x + sn * y
                Zb[:, cols[j]] = c * y - sn * x
            K = len(nd)
            dd, zz = ds[nd], zs[nd]
            orig, tau = np.zeros(K), np.zeros(K)
            for i in range(K):
                orig[i], tau[i] = root(dd, zz, rho, i)
            lamr = orig + tau
            zh = np.zeros(K)
            for j in range(K):
                p = ((orig[j] - dd[j]) + tau[j]) / rho
                for i in range(K):
                    if i != j:
                        p *= ((orig[i] - dd[j]) + tau[i]) / (dd[i] - dd[j])
                zh[j] = np.copysign(np.sqrt(abs(p)), zz[j])
            U = np.zeros((K, K))
            for i in range(K):
                u = zh / ((dd - orig[i]) - tau[i])
                U[:, i] = u / np.linalg.norm(u)
            V = Zb[:, [cols[j] for j in nd]] @ U
            dfv = np.array([dval[j] for j in df])
            dfc = [cols[j] for j in df]
            order = np.argsort(dfv, kind="stable")
            dfv, dfc = dfv[order], [dfc[q] for q in order]
            newlam, newZ = np.zeros(b - a), np.zeros((b - a, b - a))
            for i in range(K):
                p = i + np.searchsorted(dfv, lamr[i], side="left")
                newlam[p], newZ[:, p] = lamr[i], V[:, i]
            for q in range(len(dfv)):
                p = q + np.searchsorted(lamr, dfv[q], side="right")
                newlam[p], newZ[:, p] = dfv[q], Zb[:, dfc[q]]
            lamn[a:b] = newlam
            Zn[a:b, :] = 0.0
            Zn[:, a:b] = 0.0
            Zn[a:b, a:b] = newZ
        lam, Z = lamn, Zn
        s *= 2
    return lam, Z
