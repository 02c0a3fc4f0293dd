"""The divide-and-conquer oracle (oracle/tdc.py, the algorithm of
sparc_amd/csrc/tdc.hip) pinned on the CPU: eigenvalues against LAPACK and
closed forms, eigenvectors by residual and orthogonality, the device's
parallel evaluation of the deflation against dlaed2's sequential scan at
every merge, and mistakes planted in the oracle caught by those checks."""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import tdc as T


def _matrices():
    rng = np.random.default_rng(11)
    out = {}
    for n in (1, 2, 3, 17, 64):
        out[f"random{n}"] = (rng.standard_normal(n), rng.standard_normal(n - 1), None)
    n = 40
    out["clement"] = (np.zeros(n), np.sqrt(np.arange(1, n) * (n - np.arange(1, n))),
                      np.arange(-(n - 1), n, 2, dtype=float))
    out["laplacian"] = (np.full(n, 2.0), np.full(n - 1, -1.0),
                        np.sort(2.0 - 2.0 * np.cos(np.arange(1, n + 1) * np.pi / (n + 1))))
    d = rng.uniform(-1.0, 1.0, n)
    out["diagonal"] = (d, np.zeros(n - 1), np.sort(d))
    out["wilkinson"] = (np.abs(np.arange(41) - 20.0), np.ones(40), None)
    w = np.abs(np.arange(21) - 10.0)
    e = np.ones(62)
    e[20] = e[41] = 1e-12
    out["glued"] = (np.tile(w, 3), e, None)
    n = 48
    lam0 = np.repeat(np.sort(rng.uniform(-1.0, 1.0, n // 6)), 6)
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = (V * lam0) @ V.T
    _, dd, ee, _, _ = sla.lapack.dsytrd(0.5 * (A + A.T), lower=1)
    out["sixfold"] = (dd, ee, None)
    out["graded"] = (10.0 ** rng.uniform(-8, 8, 40), 10.0 ** rng.uniform(-8, 0, 39), None)
    return out


CASES = _matrices()


def _check(d, e, lam, Z, exact=None):
    n = len(d)
    Tm = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    ref = exact if exact is not None else (sla.eigh_tridiagonal(d, e, eigvals_only=True) if n > 1 else d)
    scale = max(np.max(np.abs(ref)), 1e-300)
    return (np.max(np.abs(lam - ref)) / scale, np.max(np.abs(Z.T @ Z - np.eye(n))),
            np.max(np.abs(Tm @ Z - Z * lam)) / scale)


@pytest.mark.parametrize("name", list(CASES))
def test_tdc_oracle_against_lapack_and_closed_forms(name):
    d, e, exact = CASES[name]
    lam, Z = T.tdc(d, e)
    dl, orth, res = _check(d, e, lam, Z, exact)
    assert np.all(np.diff(lam) >= 0)
    assert dl < 1e-13 and orth < 1e-13 and res < 1e-13, (dl, orth, res)


@pytest.mark.parametrize("name", ["sixfold", "wilkinson", "glued", "random64", "diagonal"])
def test_tdc_parallel_deflation_equals_scan(name):
    """At every merge the chain evaluation (the device's) makes dlaed2's
    decisions with the same values, bit for bit."""
    d, e, _ = CASES[name]
    merges = []

    def both(ds, zs, rho, tol):
        a = T.deflate_scan(ds, zs, rho, tol)
        b = T.deflate_chains(ds, zs, rho, tol)
        assert sorted(a[0]) == sorted(b[0]) and a[1] == b[1]
        assert [r[:2] for r in a[2]] == [r[:2] for r in b[2]]
        assert all(x[2] == y[2] and x[3] == y[3] for x, y in zip(a[2], b[2]))
        assert all(a[5][j] == b[5][j] for j in a[0])
        assert all(a[3][j] == b[3][j] and a[4][j] == b[4][j] for j in a[1])
        merges.append(len(a[2]))
        return b

    lam, Z = T.tdc(d, e, deflate=both)
    assert merges
    if name == "sixfold":
        assert sum(merges) > 0   # rotation chains were exercised
    assert max(_check(d, e, lam, Z)) < 1e-13


def test_tdc_planted_mistakes_are_caught():
    """A dropped rotation, deflated values placed in the wrong order and a
    secular root taken in the next pole interval each break the residual by
    far more than rounding (sixfold-degenerate spectrum: every kind of
    deflation occurs)."""
    d, e, _ = CASES["sixfold"]

    def no_rotation(*a):
        r = T.deflate_chains(*a)
        return r[0], r[1], [], r[3], r[4], r[5]

    def misplaced(*a):
        df, nd, rots, ds, zs, dval = T.deflate_chains(*a)
        vals = [dval[j] for j in df]
        return df, nd, rots, ds, zs, dict(zip(df, vals[::-1])) if len(df) > 1 else dval

    def next_interval(dd, zz, rho, i):
        return T.secular_root(dd, zz, rho, min(i + 1, len(dd) - 1))

    for kw in ({"deflate": no_rotation}, {"deflate": misplaced}, {"root": next_interval}):
        lam, Z = T.tdc(d, e, **kw)
        dl, orth, res = _check(d, e, np.sort(lam), Z[:, np.argsort(lam)])
        assert max(dl, res) > 1e-6, kw
