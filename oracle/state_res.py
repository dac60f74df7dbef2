"""TEST INFRASTRUCTURE ONLY -- the cmdata block bookkeeping of the list
builder restated over the reference's own list routines.

Add_coef (src/bsr/BSR_MAT3/add_coef.f90:2-59) sends each coefficient to the
current block of its (kpol, itype), takes a new block when the current one
holds mb records (at most kb per type, nb in all), and when none is free
generates the type with the most blocks; Gen_matrix's list step
(gen_matrix.f90:17-47) passes one block as it stands and merges several
with Merge_CDATA; State_res finally generates every (jpol, jtype)
(state_res.f90:155-159).  Initilize_cmdata (mod_cmdata.f90:102-124) gives
each type one block.  Here a block's content is made by the reference's
Add_cdata and the merge by its Merge_CDATA (both through oracle/refbind.py
on oracle/_ref/libbsr_ref.so), so the records and sums of each list are the
reference's; the block bookkeeping between them is this file's restatement.
"""
from __future__ import annotations

import numpy as np

from . import refbind as R

NTYPE = 9


def build_lists(expansion, npol, nb=2000, mb=5000, kb=100, eps_c=1e-10):
    """expansion: (itype, kpol, k1, k2, k3, k4, c) in Add_coef order.
    Returns (batch, jtype, jpol, k1, k2, k3, k4, c) in generation order."""
    itype_a, kpol_a, k1a, k2a, k3a, k4a, ca = expansion
    # chain[type] = list of blocks; a block = list of raw records (j1..j4, c)
    chain = {(k, t): [[]] for k in range(-1, npol + 1) for t in range(1, NTYPE + 1)}
    nfree = nb - len(chain)
    out = []
    nbatch = [0]

    def block_content(recs):
        if not recs:
            return [np.zeros(0, np.int32)] * 4, np.zeros(0)
        j = np.asarray(recs)
        return R.add_cdata(j[:, 0].astype(np.int32), j[:, 1].astype(np.int32), j[:, 2].astype(np.int32),
                           j[:, 3].astype(np.int32), j[:, 4])

    def count(recs):
        return len(set((int(r[0]), int(r[1]), int(r[2]), int(r[3])) for r in recs))

    def generate(t, k):
        nonlocal nfree
        ch = chain[(k, t)]
        n = len(ch)
        if n == 1 and not ch[0]:
            return
        if n == 1:
            (k1, k2, k3, k4), c = block_content(ch[0])
            for q in range(len(c)):
                out.append((nbatch[0], t, k, int(k1[q]), int(k2[q]), int(k3[q]), int(k4[q]), float(c[q])))
        else:
            K = [[], [], [], []]
            C = []
            ip, jp = [], []
            for b in ch:
                (k1, k2, k3, k4), c = block_content(b)
                ip.append(len(C) + 1)
                for q, a in enumerate((k1, k2, k3, k4)):
                    K[q].extend(int(x) for x in a)
                C.extend(float(x) for x in c)
                jp.append(len(C) if len(c) else 0)
            ipt, cc = R.merge_cdata(ip, jp, *K, np.asarray(C), eps_c)
            for i in ipt:
                i = int(i) - 1
                out.append((nbatch[0], t, k, K[0][i], K[1][i], K[2][i], K[3][i], float(cc[i])))
            nfree += n - 1
        chain[(k, t)] = [[]]
        nbatch[0] += 1

    for t, k, a, b, c4, d, c in zip(itype_a, kpol_a, k1a, k2a, k3a, k4a, ca):
        t, k = int(t), int(k)
        ch = chain[(k, t)]
        ch[-1].append((int(a), int(b), int(c4), int(d), float(c)))
        if count(ch[-1]) < mb:
            continue
        n = len(ch)
        if n < kb and nfree > 0:
            ch.append([])
            nfree -= 1
            continue
        jt, jp = t, k
        for i in range(1, NTYPE + 1):
            for kk in range(-1, npol + 1):
                if len(chain[(kk, i)]) > n:
                    jt, jp = i, kk
        generate(jt, jp)
        if (jt, jp) == (t, k):
            continue
        if nfree <= 0:
            raise RuntimeError(" Add_coef: problems with new block ")
        chain[(k, t)].append([])
        nfree -= 1
    for k in range(-1, npol + 1):
        for t in range(1, NTYPE + 1):
            generate(t, k)
    if not out:
        z = np.zeros(0, np.int32)
        return (z,) * 7 + (np.zeros(0),)
    cols = list(zip(*out))
    return tuple(np.asarray(x, np.int32) for x in cols[:7]) + (np.asarray(cols[7], np.float64),)
