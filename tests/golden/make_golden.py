"""Write the golden fixtures of this directory (run in the container that has
the reference checkout; the fixtures are committed, the reference is not).

  bs_int_out.json   R^k known answers of the reference's own BS test
                    (src/libs/BS/TEST_BS/int_out, method RK_cell) and its grid
                    (TEST_BS/knot.dat): data files of the reference's tests.
  ref_vectors.npz   outputs of the reference's Fortran routines, compiled by
                    oracle/ref.mk and driven through oracle/ref_driver.f90, on
                    small seeded inputs: gauss tables, knots of both grid
                    recipes, moments, diagonal-cell integrals, R^k bands,
                    densities, all four convolution cases, band algebra and
                    the UPDATE_* accumulators.

Usage: python tests/golden/make_golden.py
"""
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
REF = "/root/reference/src/libs/BS/TEST_BS"


def int_out():
    out = []
    head = None
    for line in open(os.path.join(REF, "int_out")):
        m = re.match(r"^R\s*(\d+)\(\s*(\w+),\s*(\w+);\s*(\w+),\s*(\w+)\)=", line)
        if m:
            head = dict(k=int(m.group(1)), orbs=[m.group(q) for q in range(2, 6)])
            continue
        if head and line.strip().endswith("RK_cell"):
            val = float(line.split()[0].replace("D", "E"))
            out.append(dict(head, rk_cell=val))
            head = None
    lines = open(os.path.join(REF, "knot.dat")).read().split("***")
    params = [float(x.split()[0]) for x in lines[0].strip().splitlines()]
    knots = [float(x) for x in lines[1].split()]
    tail = [int(x.split()[0]) for x in lines[2].strip().splitlines()[:3]]
    return dict(int_out=out, knot=dict(ks=int(params[0]), ns=int(params[1]), z=params[2], h=params[3],
                                       hmax=params[4], rmax=params[5], t5=knots, ml=tail[0], me=tail[1],
                                       nv=tail[2]))


def ref_vectors():
    from oracle import refbind as R
    out = {}
    for k in range(1, 16):
        x, w = R.gauss(k)
        out[f"gauss_x_{k}"] = x
        out[f"gauss_w_{k}"] = w
    grids = {"g0": (0, 8, 2.0, 0.125, 160.0, 160.0), "g1": (1, 8, 6.0, 0.2, 0.9, 20.0),
             "g2": (1, 4, 2.0, 0.25, 1.0, 6.0), "g3": (1, 8, 4.0, 0.25, 1.2, 7.0)}
    for name, args in grids.items():
        t, ns, ml, me, h = R.mkgrid(*args)
        out[f"{name}_args"] = np.array(args, dtype=np.float64)
        out[f"{name}_t"] = t
        out[f"{name}_ints"] = np.array([ns, ml, me])
        out[f"{name}_h"] = np.array([h])
    rng = np.random.default_rng(2024)
    for name in ("g2", "g3"):
        kind, ks, z, h0, hmax, rmax = out[f"{name}_args"]
        ks = int(ks)
        t = out[f"{name}_t"]
        ns, ml, me = (int(v) for v in out[f"{name}_ints"])
        h = float(out[f"{name}_h"][0])
        g = R.RefGrid(ks, ns, ml, me, h, t)
        gr, grw, bsp, sb = g.tables()
        out[f"{name}_gr"], out[f"{name}_grw"], out[f"{name}_bsp"], out[f"{name}_sb"] = gr, grw, bsp, sb
        p = rng.standard_normal((ns, 3))
        out[f"{name}_p"] = p
        ks_list = range(0, 4)
        for k in ks_list:
            r1, r2, rd = g.rk_moments(k)
            out[f"{name}_rkd1_{k}"], out[f"{name}_rkd2_{k}"] = r1, r2
            if ks <= 4 or k == 1:
                out[f"{name}_rkd_{k}"] = rd
            rkb = g.mrk_cell(k)
            if ks <= 4:
                out[f"{name}_rkb_{k}"] = rkb
            else:
                out[f"{name}_rkbs_{k}"] = rkb.reshape(-1, order="F")[::7].copy()
            for icase in (1, 2, 3, 4):
                typ = "s" if icase <= 2 else "x"
                d = R.density(ns, ks, p[:, 0], p[:, 1], typ)
                out[f"{name}_dens_{typ}"] = d
                out[f"{name}_conv_{k}_{icase}"] = g.convol(k, d, icase)
        ds = R.density(ns, ks, p[:, 0], p[:, 1], "s")
        dx = R.density(ns, ks, p[:, 0], p[:, 2], "s")
        out[f"{name}_sumamb_s"] = np.array([R.sum_amb(ns, ks, ds, dx, "s")])
        out[f"{name}_bav_s"] = R.bav(ns, ks, ds, p[:, 2], "s", "r")
        out[f"{name}_bxv"] = R.bxv(ks, ns, sb, p[:, 0])
        out[f"{name}_bvmv"] = np.array([R.bvmv(ns, ks, sb, p[:, 0], p[:, 1])])
        g.set_orbitals(p)
        out[f"{name}_quadr"] = np.array([[g.quadr(i, j, m) for m in (0, 1, 2, 3, -1, -2, -3)]
                                         for i in (1, 2, 3) for j in (1, 2, 3)])
        # accumulators: a fixed sequence of updates on nch=3, ncp=2
        M = R.RefMatrix(ns, 3, 2, 2)
        dfull = rng.standard_normal((ns, ns))
        dband = rng.standard_normal((ns, ks))
        v, u = rng.standard_normal(ns), rng.standard_normal(ns)
        M.update_hx(2, 1, dfull, "x")
        M.update_hx(1, 3, dfull, "x")
        M.update_hx(3, 3, dband, "s")
        M.update_hl(2, 2, 0.7)
        M.update_hl(3, 1, -1.3)
        M.update_hw(1, 2, v, u)
        M.update_hw(3, 2, u, v)
        M.update_hv(2, 1, v, 0.5)
        M.update_hb(2, 1, 0.25)
        M.update_cf(1, 3, 1, 2.0)
        hcc, hcb, hbb, acf = M.get()
        out[f"{name}_upd_in"] = np.concatenate([dfull.ravel(order="F"), dband.ravel(order="F"), v, u])
        out[f"{name}_upd_hcc"], out[f"{name}_upd_hcb"] = hcc, hcb
        out[f"{name}_upd_hbb"], out[f"{name}_upd_acf"] = hbb, acf
    return out


if __name__ == "__main__":
    with open(os.path.join(HERE, "bs_int_out.json"), "w") as f:
        json.dump(int_out(), f, indent=1)
    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **ref_vectors())
    print("fixtures written")
