"""Write tests/golden/lcore_vectors.npz: outputs of the reference's own
Gen_Lval / INT_core (src/bsr/BSR_MAT3/l_core.f90), compiled by oracle/ref.mk
and driven through oracle/ref_driver.f90, on a small seeded case with a
closed core: hl_full(:,:,l), L_int and L_vec, with every input (knots,
orbitals, l values, channel pointers, z, kclosd and the db2 band handed to
the reference).  Run in the container that has the reference checkout; the
fixture is committed, the reference is not.

Usage: python tests/golden/make_golden_lcore.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def case():
    """Knots of a 40-spline grid (ks = 8, Z = 6, Rmax = 15), seven orbitals:
    a closed core of three (l = 0, 0, 1), three bound orbitals (l = 0, 1, 2)
    and one channel orbital, random B-spline coefficients with p(1) = 0."""
    from bsr_amd import grid as G
    g = G.fixed_ns_grid(8, 40, 6.0, 15.0, 0.1)
    rng = np.random.default_rng(11)
    pbs = rng.standard_normal((g.ns, 7))
    pbs[0, :] = 0.0
    lbs = np.array([0, 0, 1, 0, 1, 2, 1], np.int32)
    iech = np.array([0, 0, 0, 0, 0, 0, 1], np.int32)
    return g, pbs, lbs, iech, 6.0, 3


def main():
    from oracle import oracle as O
    from oracle import refbind as R
    if not R.available():
        R.build()
    g, pbs, lbs, iech, z, kclosd = case()
    og = O.Grid(g.ks, g.ns, g.ml, g.me, g.h, g.t)
    db2 = og.db2()
    rg = R.RefGrid(g.ks, g.ns, g.ml, g.me, g.h, g.t)
    out = {"ks": g.ks, "ns": g.ns, "ml": g.ml, "me": g.me, "h": g.h, "t": g.t, "pbs": pbs, "lbs": lbs,
           "iech": iech, "z": z, "db2": db2}
    for kc in (0, kclosd):
        mlh, hl, li, lv = rg.gen_lval(pbs, lbs, iech, z, kc, db2)
        out[f"hl_full_k{kc}"] = hl
        out[f"lint_k{kc}"] = li
        out[f"lvec_k{kc}"] = lv
    out["kclosd"] = kclosd
    np.savez_compressed(os.path.join(HERE, "lcore_vectors.npz"), **out)


if __name__ == "__main__":
    main()
