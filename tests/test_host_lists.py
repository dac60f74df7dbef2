"""The coefficient-list builder of the C ABI (bsrg_cdata_build replacing
Add_cdata, bsrg_cdata_merge replacing Merge_CDATA; host code, no GPU)
against the reference's own routines: the fixture
tests/golden/lists_vectors.npz and, when oracle/_ref is present, the live
build on further seeded cases.  Keys and ipt are compared exactly, and the
coefficients bit for bit: the merge adds equal keys of different blocks in
the reference's order.
"""
import os

import numpy as np
import pytest

from bsr_amd import core
from oracle import refbind as R

GOLD = os.path.join(os.path.dirname(__file__), "golden", "lists_vectors.npz")


def test_build_matches_reference_fixture():
    f = np.load(GOLD)
    k, c = core.cdata_build(f["add_j1"], f["add_j2"], f["add_j3"], f["add_j4"], f["add_c"])
    for q, name in enumerate(("add_k1", "add_k2", "add_k3", "add_k4")):
        assert np.array_equal(k[q], f[name])
    assert np.array_equal(c, f["add_out"])


def test_merge_matches_reference_fixture():
    f = np.load(GOLD)
    ipt, merged = core.cdata_merge(f["mrg_ip"], f["mrg_jp"], f["mrg_k1"], f["mrg_k2"], f["mrg_k3"], f["mrg_k4"],
                                   f["mrg_c"], float(f["mrg_eps"]))
    assert np.array_equal(ipt, f["mrg_ipt"])
    assert np.array_equal(merged, f["mrg_out"])


def test_merge_properties():
    """Keys of ipt strictly increasing, each the sum of its key's entries,
    everything at or below eps dropped, empty and single blocks."""
    rng = np.random.default_rng(4)
    ks, cs, ip, jp, base = [[], [], [], []], [], [], [], 0
    for b in range(5):
        j = [rng.integers(1, 4, 80) for _ in range(4)]
        k, c = core.cdata_build(*j, rng.uniform(-1, 1, 80))
        for q in range(4):
            ks[q].append(k[q])
        cs.append(c)
        ip.append(base + 1)
        jp.append(base + len(c))
        base += len(c)
    ip.append(0)  # an empty block
    jp.append(0)
    k = [np.concatenate(x) for x in ks]
    c = np.concatenate(cs)
    ipt, merged = core.cdata_merge(np.array(ip), np.array(jp), *k, c, 0.05)
    keys = [tuple(int(k[q][i - 1]) for q in range(4)) for i in ipt]
    assert keys == sorted(set(keys))
    total = {}
    for i in range(len(c)):
        key = tuple(int(k[q][i]) for q in range(4))
        total[key] = total.get(key, 0.0) + c[i]
    for i, key in zip(ipt, keys):
        assert abs(merged[i - 1] - total[key]) <= 1e-14
    assert {kk for kk, v in total.items() if abs(v) > 0.05 + 1e-12} <= set(keys)
    assert all(abs(merged[i - 1]) > 0.05 for i in ipt)
    ipt1, _ = core.cdata_merge(np.array([1]), np.array([len(cs[0])]), *[x[: len(cs[0])] for x in k], cs[0], 0.0)
    assert list(ipt1) == list(range(1, len(cs[0]) + 1))
    ipt0, _ = core.cdata_merge(np.array([0]), np.array([0]), *k, c, 0.0)
    assert len(ipt0) == 0


@pytest.mark.skipif(not R.available(), reason="reference build oracle/_ref not present")
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_against_live_reference(seed):
    rng = np.random.default_rng(seed)
    n = 400
    j = [rng.integers(1, 5, n).astype(np.int32) for _ in range(4)]
    c = rng.uniform(-1, 1, n)
    k, co = core.cdata_build(*j, c)
    kr, cr = R.add_cdata(*j, c)
    assert all(np.array_equal(a, b) for a, b in zip(k, kr)) and np.array_equal(co, cr)
    # blocks of different lengths sharing keys; one repeats another with the
    # opposite sign so that sums cancel
    ks, cs, ip, jp, base = [[], [], [], []], [], [], [], 0
    for b in range(6):
        m = int(rng.integers(20, 90))
        jj = [rng.integers(1, 4, m).astype(np.int32) for _ in range(4)]
        cc = rng.uniform(-1, 1, m)
        if b == 4:
            jj = [x[: min(m, len(x))] for x in (ks[q][1] for q in range(4))]
            cc = -cs[1][: len(jj[0])]
        kb, cb = R.add_cdata(*jj, cc)
        for q in range(4):
            ks[q].append(kb[q])
        cs.append(cb)
        ip.append(base + 1)
        jp.append(base + len(cb))
        base += len(cb)
    kk = [np.concatenate(x) for x in ks]
    call = np.concatenate(cs)
    for eps in (0.0, 1e-10, 0.3):
        ipt, merged = core.cdata_merge(np.array(ip), np.array(jp), *kk, call, eps)
        iptr, mr = R.merge_cdata(np.array(ip, np.int32), np.array(jp, np.int32), *kk, call, eps)
        assert np.array_equal(ipt, iptr), eps
        assert np.array_equal(merged, mr), eps
