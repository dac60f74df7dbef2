"""Write tests/golden/lists_vectors.npz: the reference's own Add_cdata and
Merge_CDATA (src/bsr/BSR_MAT3/add_cdata.f90, merge_cdata.f90), compiled by
oracle/ref.mk and driven through oracle/ref_driver.f90, on seeded records:
one block built record by record (duplicates included), and a merge of
blocks that share keys, with coefficients that cancel to below eps.  Inputs
and outputs are stored; the reference is not.

Usage: python tests/golden/make_golden_lists.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def records(rng, n, span):
    return [rng.integers(1, span, n).astype(np.int32) for _ in range(4)], rng.uniform(-1, 1, n)


def blocks(rng, nblocks, per, span):
    """nblocks sorted, duplicate-free blocks stored one after another, with
    keys shared between blocks; some coefficients opposite so sums cancel."""
    from oracle import refbind as R
    ks, cs, ip, jp = [[], [], [], []], [], [], []
    base = 0
    for b in range(nblocks):
        (j1, j2, j3, j4), c = records(rng, per, span)
        if b % 3 == 2 and len(cs):  # repeat keys of the previous block with opposite sign
            m = min(per, len(cs[-1]))
            j1[:m], j2[:m], j3[:m], j4[:m] = (ks[q][-1][:m] for q in range(4))
            c[:m] = -cs[-1][:m]
        (k1, k2, k3, k4), cb = R.add_cdata(j1, j2, j3, j4, c)
        for q, k in enumerate((k1, k2, k3, k4)):
            ks[q].append(k)
        cs.append(cb)
        ip.append(base + 1)
        jp.append(base + len(cb))
        base += len(cb)
    return [np.concatenate(k) for k in ks], np.concatenate(cs), np.array(ip, np.int32), np.array(jp, np.int32)


def main():
    from oracle import refbind as R
    if not R.available():
        R.build()
    rng = np.random.default_rng(17)
    out = {}
    (j1, j2, j3, j4), c = records(rng, 600, 6)
    (k1, k2, k3, k4), co = R.add_cdata(j1, j2, j3, j4, c)
    out.update(add_j1=j1, add_j2=j2, add_j3=j3, add_j4=j4, add_c=c, add_k1=k1, add_k2=k2, add_k3=k3, add_k4=k4,
               add_out=co)
    (m1, m2, m3, m4), mc, ip, jp = blocks(rng, 7, 120, 4)
    ipt, merged = R.merge_cdata(ip.copy(), jp, m1, m2, m3, m4, mc, 1e-10)
    out.update(mrg_k1=m1, mrg_k2=m2, mrg_k3=m3, mrg_k4=m4, mrg_c=mc, mrg_ip=ip, mrg_jp=jp, mrg_ipt=ipt,
               mrg_out=merged, mrg_eps=1e-10)
    np.savez_compressed(os.path.join(HERE, "lists_vectors.npz"), **out)


if __name__ == "__main__":
    main()
