"""Knot grids for the B-spline basis.

``mkgrid`` is the reference's semi-exponential recipe (src/libs/BS/define_grid.f90:
mkgrid1 :213-276, mkgrid :86-153), evaluated by the C ABI.  ``fixed_ns_grid``
keeps that recipe's shape (ml uniform steps of h in z*r up to 1, geometric
growth by (1+h), then uniform steps of hmax up to rmax) but solves for h and
hmax so that the number of splines is exactly ns: workloads are specified by
ns, and the recipe alone does not hit an arbitrary ns.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import lib as L


@dataclass
class Grid:
    ks: int
    ns: int
    ml: int
    me: int
    h: float
    t: np.ndarray
    z: float
    rmax: float
    hmax: float

    @property
    def nv(self):
        return self.ns - self.ks + 1


def mkgrid(kind: int, ks: int, z: float, h: float, hmax: float, rmax: float):
    cap = int((rmax * z / h + 1 + ks) * 2) + 4 * ks + 16
    t = np.zeros(cap, order="F")
    ns, ml, me, hh = C.c_int(), C.c_int(), C.c_int(), C.c_double()
    L.check(L.lib().bsrg_mkgrid(kind, ks, z, h, hmax, rmax, cap, t, ns, ml, me, hh))
    return t[: ns.value + ks].copy(), ns.value, ml.value, me.value, hh.value


def recipe_grid(kind: int, ks: int, z: float, h: float, hmax: float, rmax: float) -> Grid:
    """The grid the reference's recipe itself gives (kind 0 mkgrid, 1
    mkgrid1), e.g. the knot.dat parameters of a BSR example."""
    t, ns, ml, me, hh = mkgrid(kind, ks, z, h, hmax, rmax)
    return Grid(ks, ns, ml, me, hh, t, z, rmax, hmax)


def fixed_ns_grid(ks: int, ns: int, z: float, rmax: float, h: float = 0.05) -> Grid:
    """mkgrid1-shaped grid with exactly ns splines.  Tries ml = round(1/h)
    first (the requested near-origin step), then coarser ml, bisecting on
    hmax (the spline count falls as hmax grows)."""
    ml0 = max(1, int(round(1.0 / h)))
    for ml in range(ml0, 0, -1):
        hh = 1.0 / ml
        lo, hi = hh / z, rmax
        n_lo = mkgrid(1, ks, z, hh, lo, rmax)[1]
        n_hi = mkgrid(1, ks, z, hh, hi, rmax)[1]
        if not (n_hi <= ns <= n_lo):
            continue
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            n_mid = mkgrid(1, ks, z, hh, mid, rmax)[1]
            if n_mid == ns:
                t, n, mlo, me, hout = mkgrid(1, ks, z, hh, mid, rmax)
                return Grid(ks, n, mlo, me, hout, t, z, rmax, mid)
            if n_mid > ns:
                lo = mid
            else:
                hi = mid
            if hi - lo < 1e-13 * hi:
                break
    raise ValueError(f"no mkgrid1-shaped grid with ns={ns} (ks={ks}, z={z}, rmax={rmax})")


def tables(g: Grid):
    """(gr, grw, bsp, sb) of the grid: gr,grw (nv,ks); bsp (nv,ks,ks) with
    bsp[iv,m,ith] = B_{iv+ith}(gr[iv,m]); sb (ns,ks) lower band."""
    nv, ks, ns = g.nv, g.ks, g.ns
    gr = np.zeros(nv * ks, order="F")
    grw = np.zeros(nv * ks, order="F")
    bsp = np.zeros(nv * ks * ks, order="F")
    sb = np.zeros((ns, ks), order="F")
    L.check(L.lib().bsrg_grid_tables(ks, ns, L.f64(g.t), gr, grw, bsp, sb))
    return gr.reshape(nv, ks), grw.reshape(nv, ks), bsp.reshape(nv, ks, ks), sb


def band_to_dense_lower(sb: np.ndarray) -> np.ndarray:
    """Full symmetric matrix from a lower band (ns,ks): (i,jp) -> (i, i+jp-ks)."""
    ns, ks = sb.shape
    S = np.zeros((ns, ns))
    for jp in range(ks):
        off = jp - (ks - 1)
        i = np.arange(ns)
        j = i + off
        ok = j >= 0
        S[i[ok], j[ok]] = sb[i[ok], jp]
        S[j[ok], i[ok]] = sb[i[ok], jp]
    return S
