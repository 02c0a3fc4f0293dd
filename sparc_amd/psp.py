"""psp8 (ONCVPSP / ABINIT format 8) pseudopotential reader and the cubic
spline SPARC uses to put radial tables on the grid.

Reader: read_pseudopotential_PSP, SPARC src/readfiles.c:1957-2324.  Stored as
the reference stores them (PSD_OBJ, src/include/isddft.h:126-150):
  * UdV = u_lp(r) / r per projector, with UdV(0) := UdV(dr)      (:2158-2162)
  * Gamma = the KB energies of the block header lines             (:2150)
  * rVloc = r * Vloc(r), Vloc_0 = Vloc(0)                         (:2169-2176)
  * rho_c_table = model core charge / 4 pi (NLCC), zero if fchrg=0 (:2226-2234)
  * rhoIsoAtom  = isolated-atom valence density / 4 pi            (:2237-2242)
  * rc[l] extended to the last r with |UdV| > 1e-8 beyond r_core   (:2281-2298)

Spline: getYD_gen (not-a-knot end conditions in derivative form) and
SplineInterpUniform (cubic Hermite per interval), src/tools.c:372-603.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

PSP_DIR = Path(__file__).resolve().parent / "psps"


def spline_derivs(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """Knot derivatives of the not-a-knot cubic spline through (x, y): the
    tridiagonal system of getYD_gen solved by forward elimination and back
    substitution (src/tools.c:372-439)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    n = x.size
    A = np.zeros(n); B = np.zeros(n); C = np.zeros(n); D = np.zeros(n)
    h = np.diff(x)
    r = np.diff(y) / h
    B[0] = h[1] * (h[0] + h[1])
    C[0] = (h[0] + h[1]) ** 2
    D[0] = r[0] * (3 * h[0] * h[1] + 2 * h[1] * h[1]) + r[1] * h[0] * h[0]
    h0, h1, r0, r1 = h[:-1], h[1:], r[:-1], r[1:]
    A[1:-1] = h1
    B[1:-1] = 2 * (h0 + h1)
    C[1:-1] = h0
    D[1:-1] = 3 * (r0 * h1 + r1 * h0)
    h0, h1, r0, r1 = h[-2], h[-1], r[-2], r[-1]
    A[-1] = (h0 + h1) ** 2
    B[-1] = h0 * (h0 + h1)
    D[-1] = r0 * h1 * h1 + r1 * (3 * h0 * h1 + 2 * h0 * h0)
    F = np.zeros(n)
    b = B[0]
    D[0] /= b
    for i in range(1, n):
        F[i] = C[i - 1] / b
        b = B[i] - A[i] * F[i]
        if abs(b) < 1e-50:
            raise ZeroDivisionError("spline_derivs: singular tridiagonal system")
        D[i] = (D[i] - D[i - 1] * A[i]) / b
    for i in range(n - 2, -1, -1):
        D[i] -= D[i + 1] * F[i + 1]
    return D


def spline_eval_uniform(x: np.ndarray, y: np.ndarray, yd: np.ndarray, xq: np.ndarray,
                        fill: float = 0.0) -> np.ndarray:
    """Cubic Hermite evaluation on a uniform knot grid (SplineInterpUniform,
    src/tools.c:552-603).  Points beyond the last knot get ``fill`` (the
    reference leaves them unset and every caller overwrites them)."""
    xq = np.asarray(xq, np.float64)
    out = np.full(xq.shape, fill, np.float64)
    dx1 = x[1] - x[0]
    inside = xq <= x[-1]
    q = xq[inside]
    j = np.floor((q - x[0]) / dx1).astype(np.int64)
    j = np.minimum(j, x.size - 2)
    p1, p3 = x[j], x[j + 1]
    dxi = 1.0 / (p3 - p1)
    dy = (y[j + 1] - y[j]) * dxi
    a0, a1 = y[j], yd[j]
    a2 = dxi * (3.0 * dy - 2.0 * yd[j] - yd[j + 1])
    a3 = dxi * dxi * (-2.0 * dy + yd[j] + yd[j + 1])
    t = q - p1
    out[inside] = ((a3 * t + a2) * t + a1) * t + a0
    return out


@dataclass
class Psp8:
    path: str
    element: str
    zatom: float
    zion: float          # Znucl: valence charge
    pspxc: int
    lmax: int
    lloc: int
    size: int
    fchrg: float
    rc: np.ndarray       # per l, extended
    ppl: np.ndarray      # projectors per l
    r: np.ndarray
    UdV: np.ndarray      # [nproj_radial, size]
    Gamma: np.ndarray    # [nproj_radial]
    rVloc: np.ndarray
    Vloc_0: float
    rho_c_table: np.ndarray
    rhoIsoAtom: np.ndarray
    is_r_uniform: bool
    spl: dict = field(default_factory=dict)

    @property
    def rchrg(self) -> float:
        """Last radial grid point (the reference's rchrg in the setup code)."""
        return float(self.r[-1])

    @property
    def rc_max(self) -> float:
        return float(max(self.rc[l] for l in range(self.lmax + 1)))

    @property
    def nproj_total(self) -> int:
        """Projectors per atom including the 2l+1 m-components."""
        return int(sum(self.ppl[l] * (2 * l + 1) for l in range(self.lmax + 1) if l != self.lloc))

    def spline(self, name: str, k: int | None = None):
        key = (name, k)
        if key not in self.spl:
            tab = getattr(self, name) if k is None else getattr(self, name)[k]
            self.spl[key] = spline_derivs(self.r, tab)
        return self.spl[key]

    def eval(self, name: str, rq: np.ndarray, k: int | None = None) -> np.ndarray:
        tab = getattr(self, name) if k is None else getattr(self, name)[k]
        if not self.is_r_uniform:
            raise NotImplementedError("non-uniform psp8 radial grids are not supported")
        return spline_eval_uniform(self.r, tab, self.spline(name, k), rq)


def read_psp8(path: str | Path) -> Psp8:
    path = Path(path)
    if not path.is_absolute() and not path.exists():
        path = PSP_DIR / path
    lines = path.read_text().splitlines()
    head = lines[0].split()
    element = head[0]
    i_rc = head.index("r_core=")
    zatom, zion = (float(v) for v in lines[1].split()[:2])
    t = lines[2].split()
    pspcod, pspxc, lmax, lloc, size = (int(v) for v in t[:5])
    if pspcod != 8:
        raise ValueError(f"{path}: pspcod {pspcod} is not 8")
    rchrg_, fchrg, qchrg = (float(v) for v in lines[3].split()[:3])
    ppl = np.array([int(v) for v in lines[4].split()[:lmax + 1]])
    ext = int(lines[5].split()[0])
    if ext in (2, 3):
        raise NotImplementedError(f"{path}: spin-orbit psp8 files are not supported")
    rc = np.array([float(v) for v in head[i_rc + 1:i_rc + 1 + lmax + 1]])
    lpos = np.concatenate([[0], np.cumsum(ppl)])
    nproj = int(lpos[-1])
    r = np.zeros(size)
    UdV = np.zeros((nproj, size))
    Gamma = np.zeros(nproj)
    rVloc = np.zeros(size)
    Vloc_0 = 0.0
    ln = 6

    def block(ncol):
        nonlocal ln
        a = np.array([[float(v) for v in lines[ln + j].split()[:ncol]] for j in range(size)])
        ln += size
        return a

    for l in range(lmax + 1):
        hdr = lines[ln].split()
        ln += 1
        if l != lloc:
            Gamma[lpos[l]:lpos[l + 1]] = [float(v) for v in hdr[1:1 + ppl[l]]]
            a = block(2 + ppl[l])
            r = a[:, 1].copy()
            with np.errstate(divide="ignore", invalid="ignore"):
                UdV[lpos[l]:lpos[l + 1]] = (a[:, 2:2 + ppl[l]] / r[:, None]).T
            UdV[lpos[l]:lpos[l + 1], 0] = UdV[lpos[l]:lpos[l + 1], 1]
        else:
            a = block(3)
            r = a[:, 1].copy()
            rVloc = r * a[:, 2]
            Vloc_0 = float(a[0, 2])
    if lloc > lmax:
        hdr = lines[ln].split()
        ln += 1
        if int(hdr[0]) != 4:
            raise ValueError(f"{path}: expected the local potential block (l = 4)")
        a = block(3)
        r = a[:, 1].copy()
        rVloc = r * a[:, 2]
        Vloc_0 = float(a[0, 2])
    rho_c = np.zeros(size)
    if fchrg > 1e-12:
        rho_c = block(3)[:, 2] / (4.0 * np.pi)
    rho_iso = block(3)[:, 2] / (4.0 * np.pi)
    dr = np.diff(r)
    if np.any(np.abs(dr) < 1e-12):
        raise ValueError(f"{path}: repeated radial grid point")
    uniform = bool(np.all(np.abs(dr - dr[0]) <= 1e-12))
    rc = rc.copy()
    for l in range(lmax + 1):
        if l == lloc:
            continue
        rc0 = rc[l]
        for k in range(lpos[l], lpos[l + 1]):
            big = (r > rc0) & (np.abs(UdV[k]) > 1e-8)
            if big.any():
                rc[l] = max(rc[l], float(r[big].max()))
    return Psp8(str(path), element, zatom, zion, pspxc, lmax, lloc, size, fchrg, rc, ppl, r, UdV,
                Gamma, rVloc, Vloc_0, rho_c, rho_iso, uniform)
