"""Write dbs_vectors.npz: outputs of the reference's Dirac B-spline routines
(src/libs/DBS, compiled by oracle/ref.mk into oracle/_ref/libdbs_ref.so and
driven through oracle/ref_dbs_driver.f90), run in the container that has the
reference checkout; the fixture is committed, the reference is not.

Grids: the reference's own DBS test grid (src/libs/DBS/test_int/knot.dat:
ks=9, ns=78, P splines of order ksp=8, Q splines of order ksq=9) and a small
seeded geometric grid with ksp=7, ksq=8.  Stored per grid: knots, Gauss
points r and weights of the intervals, P and Q B-spline values, moments of
both bases for several powers, the R^k bands of the four families
(pppp, qqqq, pqpq, qpqp; every 31st element for the small grid, every 197th
for the test grid), and the four convolution cases with seeded densities
(a rank-1 exchange density; exchange results every 7th element).

Usage: python tests/golden/make_golden_dbs.py
"""
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
KNOT = "/root/reference/src/libs/DBS/test_int/knot.dat"


def test_grid():
    txt = open(KNOT).read()
    body = txt.split("grid points:")[1].split("***")[0]
    pts = re.findall(r"^\s*(\d+)\s+([-0-9.D+E]+)\s*$", body, re.M)
    return np.array([float(v.replace("D", "E")) for _, v in pts])


def main():
    from oracle import oracle as O
    from oracle import refbind as R
    R.build()
    out = {}
    t_small = np.concatenate([np.zeros(8), np.geomspace(0.02, 10.0, 30), np.full(7, 10.0)])
    grids = {"d1": (9, 78, 8, 9, test_grid(), 197), "d2": (8, len(t_small) - 8, 7, 8, t_small, 31)}
    rng = np.random.default_rng(2024)
    for name, (ks, ns, ksp, ksq, t, stride) in grids.items():
        r = R.RefDbs(ks, ns, ksp, ksq, t)
        out[f"{name}_ints"] = np.array([ks, ns, ksp, ksq])
        out[f"{name}_t"] = t
        tb = r.tables()
        nv = ns - ks + 1
        out[f"{name}_gr"] = tb["gr"]
        out[f"{name}_grw"] = tb["grw"]
        out[f"{name}_pbsp"] = tb["pbsp"][:nv]
        out[f"{name}_qbsp"] = tb["qbsp"][:nv]
        for q in (0, 1):
            kb = ksq if q else ksp
            for k in (0, 3, -1, -5):
                out[f"{name}_mom_{q}_{k}"] = r.moments(q, k)[: kb * (kb + 1) // 2]
        for f in range(4):
            for k in (0, 3):
                rkb = r.mrk(f, k)
                out[f"{name}_rkb_{f}_{k}"] = rkb.reshape(-1, order="F")[::stride]
        d_s = np.asfortranarray(rng.standard_normal((ns, ks)))
        u, v = rng.standard_normal(ns), rng.standard_normal(ns)
        d_x = np.asfortranarray(np.outer(u, v))
        out[f"{name}_dens_s"] = d_s
        out[f"{name}_dens_x_uv"] = np.stack([u, v])
        for f in range(4):
            for icase in (1, 2, 3, 4):
                a = r.convol(f, 2, icase, d_s if icase <= 2 else d_x)
                out[f"{name}_conv_{f}_{icase}"] = a if icase <= 2 else a.reshape(-1, order="F")[::7]
    np.savez_compressed(os.path.join(HERE, "dbs_vectors.npz"), **out)
    print("wrote", os.path.join(HERE, "dbs_vectors.npz"))


if __name__ == "__main__":
    main()
