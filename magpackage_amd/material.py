"""B-H materials: the reference's ``MaterialLib`` tables and their packing.

- ``create_linear``      src/MaterialLib.m:58-63
- ``create_nonlinear``   src/MaterialLib.m:65-125: sort / unique('last') the
  B-H data, nu = H/B (nu(0) = nu(2)), append the vacuum-guide knot at
  B^2_end + 0.1, PCHIP in B^2 space, MaxBSq = last B^2.
- ``pchip``              MATLAB's ``pchip`` (Fritsch-Carlson slopes with the
  harmonic-mean interior rule and the three-point one-sided end rule, as
  published by Moler, "Numerical Computing with MATLAB", ch. 3) returning
  a pp-form with coefficients ``[c3 c2 c1 c0]`` per piece (``pwch``).
- ``pack_material_data`` src/Assembler.m:270-304: the 8 SoA arrays the
  Jacobian kernel reads.
- ``marrocco``           tabulates nu(s) = (k1 exp(k2 s) + k3)/mu0 (s = B^2)
  into the same pp-form as cubic Hermite pieces on uniform s-knots.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

MU0 = 4.0 * np.pi * 1e-7


@dataclass
class PP:
    breaks: np.ndarray  # (n+1,)
    coefs: np.ndarray   # (n, 4) descending powers

    @property
    def pieces(self) -> int:
        return self.coefs.shape[0]


@dataclass
class MaterialData:
    Type: str
    nu0: float = 1.0 / MU0
    Nu_Linear: float = 0.0
    MaxBSq: float | None = None
    SplineNu: PP | None = None
    extra: dict = field(default_factory=dict)


def create_linear(mu_r: float) -> MaterialData:
    return MaterialData(Type="Linear", Nu_Linear=1.0 / (mu_r * MU0), nu0=1.0 / MU0)


def _sign(x):
    return np.sign(x)


def _pchip_end(h1, h2, del1, del2):
    d = ((2.0 * h1 + h2) * del1 - h1 * del2) / (h1 + h2)
    if _sign(d) != _sign(del1):
        d = 0.0
    elif _sign(del1) != _sign(del2) and abs(d) > abs(3.0 * del1):
        d = 3.0 * del1
    return d


def pchip_slopes(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    n = x.size
    h = np.diff(x)
    dl = np.diff(y) / h
    d = np.zeros(n)
    if n == 2:
        d[:] = dl[0]
        return d
    for k in range(n - 2):
        if _sign(dl[k]) * _sign(dl[k + 1]) > 0:
            h1, h2 = h[k], h[k + 1]
            hs = h1 + h2
            w1 = (h1 + hs) / (3.0 * hs)
            w2 = (hs + h2) / (3.0 * hs)
            dmax = max(abs(dl[k]), abs(dl[k + 1]))
            dmin = min(abs(dl[k]), abs(dl[k + 1]))
            d[k + 1] = dmin / (w1 * (dl[k] / dmax) + w2 * (dl[k + 1] / dmax))
    d[0] = _pchip_end(h[0], h[1], dl[0], dl[1])
    d[n - 1] = _pchip_end(h[n - 2], h[n - 3], dl[n - 2], dl[n - 3])
    return d


def hermite_pp(x: np.ndarray, y: np.ndarray, s: np.ndarray) -> PP:
    """pp-form of the cubic Hermite interpolant with slopes ``s`` (MATLAB pwch)."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    s = np.asarray(s, dtype=np.float64)
    dx = np.diff(x)
    divdif = np.diff(y) / dx
    dzzdx = (divdif - s[:-1]) / dx
    dzdxdx = (s[1:] - divdif) / dx
    coefs = np.stack([(dzdxdx - dzzdx) / dx, 2.0 * dzzdx - dzdxdx, s[:-1], y[:-1]], axis=1)
    return PP(breaks=x.copy(), coefs=coefs)


def pchip(x, y) -> PP:
    x = np.asarray(x, dtype=np.float64).ravel()
    y = np.asarray(y, dtype=np.float64).ravel()
    return hermite_pp(x, y, pchip_slopes(x, y))


def ppval(pp: PP, s):
    s = np.asarray(s, dtype=np.float64)
    k = np.clip(np.searchsorted(pp.breaks, s, side="right") - 1, 0, pp.pieces - 1)
    dx = s - pp.breaks[k]
    c = pp.coefs[k]
    return ((c[..., 0] * dx + c[..., 1]) * dx + c[..., 2]) * dx + c[..., 3]


def create_nonlinear(B_curve, H_curve) -> MaterialData:
    nu0 = 1.0 / MU0
    B = np.abs(np.asarray(B_curve, dtype=np.float64).ravel())
    H = np.abs(np.asarray(H_curve, dtype=np.float64).ravel())
    order = np.argsort(B, kind="stable")
    B, H = B[order], H[order]
    # unique(B, 'last'): keep the last occurrence of each value
    last = np.r_[B[1:] != B[:-1], True]
    Bu, Hu = B[last], H[last]
    Bsq = Bu ** 2
    Nu = np.zeros_like(Bu)
    nz = Bu > 1e-12
    Nu[nz] = Hu[nz] / Bu[nz]
    if not nz[0]:
        Nu[0] = Nu[1] if Bu.size > 1 else nu0
    if Bsq.size > 2:
        B_end, Nu_end = Bu[-1], Nu[-1]
        slope_vacuum = (nu0 - Nu_end) / (2.0 * B_end ** 2)
        dBsq = 0.1
        Nu_new = Nu_end + slope_vacuum * dBsq
        if Nu_new > nu0:
            Nu_new = nu0
        Bsq = np.r_[Bsq, Bsq[-1] + dBsq]
        Nu = np.r_[Nu, Nu_new]
    if Bsq.size > 1:
        spl = pchip(Bsq, Nu)
    else:
        spl = PP(breaks=Bsq.copy(), coefs=np.array([[0.0, 0.0, 0.0, Nu[0]]]))
    return MaterialData(Type="Nonlinear", nu0=nu0, MaxBSq=float(Bsq.max()), SplineNu=spl,
                        extra={"B_data": Bu, "H_data": Hu})


def marrocco(k1: float = 5e-4, k2: float = 5.0, k3: float = 2e-4, s_max: float = 4.84,
             pieces: int = 44) -> MaterialData:
    """Marrocco reluctivity tabulated to the reference pp format.

    nu(s) = (k1 exp(k2 s) + k3) / mu0 on uniform s-knots over [0, s_max],
    cubic Hermite pieces from the exact values and derivatives.  Clamping to
    nu0 above the vacuum value is left to the kernel (MexMaterialUtils.hpp:80-86).
    """
    s = np.linspace(0.0, s_max, pieces + 1)
    nu = (k1 * np.exp(k2 * s) + k3) / MU0
    dnu = (k1 * k2 * np.exp(k2 * s)) / MU0
    return MaterialData(Type="Nonlinear", nu0=1.0 / MU0, MaxBSq=float(s_max),
                        SplineNu=hermite_pp(s, nu, dnu),
                        extra={"k1": k1, "k2": k2, "k3": k3})


def marrocco_exact(s, k1=5e-4, k2=5.0, k3=2e-4):
    """Exact Marrocco nu(s) and dnu/ds (for tests of the tabulation)."""
    s = np.asarray(s, dtype=np.float64)
    return (k1 * np.exp(k2 * s) + k3) / MU0, (k1 * k2 * np.exp(k2 * s)) / MU0


def pack_material_data(mat_lib) -> dict:
    """The 8 packed arrays of Assembler.packMaterialData (Assembler.m:270-304).

    ``mat_lib``: dict tag -> MaterialData (a containers.Map) or a sequence
    whose i-th entry is tag i+1.  Arrays are sized maxTag+1 and indexed by
    tag-1 on the device, as in the reference.
    """
    if isinstance(mat_lib, dict):
        keys = sorted(int(k) for k in mat_lib)
        get = lambda t: mat_lib[t]
        max_tag = max(keys)
    else:
        keys = list(range(1, len(mat_lib) + 1))
        get = lambda t: mat_lib[t - 1]
        max_tag = len(mat_lib)
    n = max_tag + 1
    lin = np.zeros(n); isn = np.zeros(n); maxb = np.zeros(n)
    starts = np.zeros(n); counts = np.zeros(n); coe_starts = np.zeros(n)
    brk, coe = [], []
    bptr = 0
    cptr = 0
    for tag in keys:
        if tag <= 0:
            continue
        m = get(tag)
        if m.Type == "Linear":
            lin[tag - 1] = m.Nu_Linear
            isn[tag - 1] = 0.0
        else:
            lin[tag - 1] = m.nu0
            isn[tag - 1] = 1.0
            maxb[tag - 1] = m.MaxBSq if m.MaxBSq is not None else 1e10
            pp = m.SplineNu
            b = pp.breaks.ravel()
            c = pp.coefs.ravel()  # row-major (piece, power) == (c_mat')(:)
            starts[tag - 1] = bptr
            counts[tag - 1] = b.size - 1
            coe_starts[tag - 1] = cptr
            brk.append(b)
            coe.append(c)
            bptr += b.size
            cptr += c.size
    # MATLAB 1-based arrays linNu(tag) -> C index tag-1: the packed arrays are
    # stored so that element [tag-1] holds tag's data (length maxTag+1).
    return {
        "linNu": lin, "isNon": isn, "maxB": maxb,
        "breaks": np.concatenate(brk) if brk else np.zeros(0),
        "coefs": np.concatenate(coe) if coe else np.zeros(0),
        "starts": starts, "counts": counts, "coe_starts": coe_starts,
    }
