"""Pins of the Dirac R^k assembly of the CPU oracle (oracle/dbsr_oracle.c,
R_data over the four families with the UPDATE_* accumulators), no GPU:

  * golden vectors written by the reference's own R_data / updates /
    Target_h (tests/golden/dirac_vectors.npz, make_golden_dirac.py) on two
    waves covering every structure of itype 1..4 (HB, HV, HW, Target_h,
    overlaps with and without perturbers, direct, exchange)
  * the live reference build (oracle/_ref/libdbs_ref.so) on larger waves
    (P order 8, Q order 9; equal orders), when it is present
  * linearity in the coefficients (doubling every c doubles every output
    exactly) and the data distinguishing the plausible slips: the PQ and QP
    quadrants of a channel block are not transposes of each other, so a
    transposed or family-swapped placement would fail the vectors above
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from oracle import refbind as R

GOLD = os.path.join(os.path.dirname(__file__), "golden", "dirac_vectors.npz")


def _run_oracle(ints, t, pq, ipbs, ich, jch, vch, iptar, lists, c):
    ks, ns, ksp, ksq, nch, npert, mk, ntarg = (int(v) for v in ints)
    g = O.DbsGrid(ks, ns, ksp, ksq, t)
    o = O.DbsrWave(g, pq, ipbs, ich, jch, vch, nch, npert, mk, ntarg, iptar)
    it, kp = lists[0], lists[1]
    r = 0
    while r < len(c):
        e = r + 1
        while e < len(c) and it[e] == it[r] and kp[e] == kp[r]:
            e += 1
        o.rdata(int(it[r]), int(kp[r]), *(lists[q][r:e] for q in range(2, 6)), c[r:e])
        r = e
    out = o.get()
    out["_g"] = g
    return out


def _gold_inputs(gv, tag):
    return [gv[f"{tag}_{k}"] for k in ("ints", "t", "pq", "ipbs", "ich", "jch", "vch", "iptar", "lists", "c")]


@pytest.fixture(scope="module")
def gv():
    return np.load(GOLD)


@pytest.mark.parametrize("tag", ["p", "n"])
def test_oracle_vs_golden(gv, tag):
    out = _run_oracle(*_gold_inputs(gv, tag))
    got = out["hch"].reshape(-1, order="F")[::5]
    assert np.array_equal(got, gv[f"{tag}_hch5"])
    for key in ("hcp", "hp", "acf", "htarg"):
        assert out[key].shape == gv[f"{tag}_{key}"].shape, key
        assert np.array_equal(out[key], gv[f"{tag}_{key}"]), key


def test_golden_covers_every_structure(gv):
    lists, c = gv["p_lists"], gv["p_c"]
    it, k3, k4 = lists[0], lists[4], lists[5]
    s1 = it == 1
    assert np.any(s1 & (k3 < 0) & (k4 < 0))      # HB
    assert np.any(s1 & (k3 > 0) & (k4 < 0))      # HV
    assert np.any(s1 & (k3 > 0) & (k4 > 0))      # HW
    assert np.any(s1 & (k3 > 0) & (k4 == 0))     # Target_h
    s2 = it == 2
    assert np.any(s2 & (k4 > 0)) and np.any(s2 & (k4 < 0))
    assert np.any(it == 3) and np.any(it == 4)
    for key in ("hcp", "hp", "acf", "htarg"):
        assert np.any(gv[f"p_{key}"] != 0.0), key


def test_quadrants_distinguish_transposes(gv):
    ints = gv["p_ints"]
    ns, nch = int(ints[1]), int(ints[4])
    out = _run_oracle(*_gold_inputs(gv, "p"))
    h = out["hch"]
    off = h[:, :, nch:]   # off-diagonal channel blocks
    pq, qp = off[:ns, ns:, :], off[ns:, :ns, :]
    assert np.max(np.abs(pq - np.transpose(qp, (1, 0, 2)))) > 1e-6 * np.max(np.abs(off))
    pp, qq = off[:ns, :ns, :], off[ns:, ns:, :]
    assert np.max(np.abs(pp - qq)) > 1e-6 * np.max(np.abs(off))


def test_linearity(gv):
    inp = _gold_inputs(gv, "p")
    a = _run_oracle(*inp)
    inp2 = list(inp)
    inp2[-1] = 2.0 * inp[-1]
    b = _run_oracle(*inp2)
    for key in ("hch", "hcp", "hp", "acf", "htarg"):
        assert np.array_equal(b[key], 2.0 * a[key]), key


@pytest.mark.skipif(not R.dbs_available(), reason="reference DBS build not present")
@pytest.mark.parametrize("shape", [(21, 9, 40, 8, 9, 4, 6, 2, 4, 0.15), (22, 9, 36, 9, 9, 5, 4, 0, 3, 0.2)])
def test_oracle_vs_live_reference(shape):
    from bsr_amd import synth_dirac as SD
    seed, ks, ns, ksp, ksq, nb, nch, npert, mk, fill = shape
    w = SD.make_dirac_wave(seed, ks, ns, ksp, ksq, 54.0, 30.0, nb, nch, npert, mk, fill)
    o = O.dirac_assemble(w)
    r = R.dirac_assemble(w)
    for key in ("hch", "hcp", "hp", "acf", "htarg"):
        assert np.array_equal(o[key], r[key]), key
