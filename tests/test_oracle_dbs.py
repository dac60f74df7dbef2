"""Pins of the Dirac B-spline (DBS) part of the CPU oracle (no GPU).

  * golden vectors written by the reference's own DBS routines
    (tests/golden/dbs_vectors.npz, see make_golden_dbs.py) on the
    reference's DBS test grid (ksp=8, ksq=9) and a small grid (ksp=7, ksq=8)
  * identities: with ksp = ksq the four families coincide; PQPQ and QPQP are
    the same integrals with the electrons swapped; bands are non-negative
  * the BS band of the same grid (the scaling law switched off): the PPPP band
    with ksp = ks is the BS R^k band, up to the two Gauss-Legendre tables
  * the live reference build, when oracle/_ref/libdbs_ref.so is present
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from oracle import refbind as R

GOLD = os.path.join(os.path.dirname(__file__), "golden", "dbs_vectors.npz")


@pytest.fixture(scope="module")
def gv():
    return np.load(GOLD)


def _grid(gv, name):
    ks, ns, ksp, ksq = (int(v) for v in gv[f"{name}_ints"])
    return O.DbsGrid(ks, ns, ksp, ksq, gv[f"{name}_t"])


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("name", ["d1", "d2"])
def test_tables_vs_reference(gv, name):
    g = _grid(gv, name)
    tb = g.tables()
    assert rel(tb["gr"], gv[f"{name}_gr"]) == 0.0
    assert rel(tb["grw"], gv[f"{name}_grw"]) == 0.0
    nv = g.nv
    assert rel(tb["pbsp"][:nv], gv[f"{name}_pbsp"]) < 1e-15
    assert rel(tb["qbsp"][:nv], gv[f"{name}_qbsp"]) < 1e-15
    # partition of unity of both bases at every Gauss point
    assert np.allclose(tb["pbsp"][:nv].sum(axis=2), 1.0, atol=1e-14)
    assert np.allclose(tb["qbsp"][:nv].sum(axis=2), 1.0, atol=1e-14)


@pytest.mark.parametrize("name", ["d1", "d2"])
def test_moments_vs_reference(gv, name):
    g = _grid(gv, name)
    for q in (0, 1):
        kb = g.ksq if q else g.ksp
        for k in (0, 3, -1, -5):
            got = g.moments(q, k)[: kb * (kb + 1) // 2]
            assert rel(got, gv[f"{name}_mom_{q}_{k}"]) < 1e-14, (q, k)


@pytest.mark.parametrize("name", ["d1", "d2"])
def test_bands_vs_reference(gv, name):
    g = _grid(gv, name)
    stride = 197 if name == "d1" else 31
    for f in range(4):
        for k in (0, 3):
            got = g.mrk(f, k).reshape(-1, order="F")[::stride]
            ref = gv[f"{name}_rkb_{f}_{k}"]
            assert np.array_equal(got != 0, ref != 0), (f, k)
            m = ref != 0
            assert np.max(np.abs(got[m] - ref[m]) / ref[m]) < 1e-14, (f, k)


@pytest.mark.parametrize("name", ["d1", "d2"])
def test_convolutions_vs_reference(gv, name):
    g = _grid(gv, name)
    ns, ks = g.ns, g.ks
    d_s = gv[f"{name}_dens_s"]
    u, v = gv[f"{name}_dens_x_uv"]
    d_x = np.asfortranarray(np.outer(u, v))
    for f in range(4):
        rkb = g.mrk(f, 2)
        for icase in (1, 2, 3, 4):
            got = O.convol(ns, ks, d_s if icase <= 2 else d_x, icase, rkb)
            ref = gv[f"{name}_conv_{f}_{icase}"]
            if icase > 2:
                got = got.reshape(-1, order="F")[::7]
            assert rel(got, ref) < 1e-13, (f, icase)


def test_family_identities():
    t = np.concatenate([np.zeros(7), np.geomspace(0.05, 8.0, 22), np.full(6, 8.0)])
    ks, ns = 7, len(t) - 7
    same = O.DbsGrid(ks, ns, ks, ks, t)
    b0 = same.mrk(0, 2)
    for f in (1, 2, 3):
        assert np.array_equal(same.mrk(f, 2), b0)
    mixed = O.DbsGrid(ks, ns, ks - 1, ks, t)
    pq, qp = mixed.mrk(2, 3), mixed.mrk(3, 3)
    # electrons swapped: rkb_pqpq(i,j,ip,jp) = rkb_qpqp(j,i,jp,ip)
    assert rel(pq, np.transpose(qp, (1, 0, 3, 2))) < 1e-14
    for f in range(4):
        assert np.all(mixed.mrk(f, 1) >= 0.0)


def test_pppp_equals_bs_band():
    """Same knots and orders: the Dirac PPPP band is the BS band (no scaling
    law on either side); only the Gauss-Legendre tables differ (Newton vs the
    reference's 15-digit table)."""
    t, ns, ml, me, h = O.mkgrid1(8, 4.0, 0.25, 1.2, 7.0)
    bs = O.Grid(8, ns, ml, 0, h, t)
    dbs = O.DbsGrid(8, ns, 8, 8, t)
    for k in (0, 2, 5):
        assert rel(dbs.mrk(0, k), bs.mrk_cell(k)) < 1e-13


@pytest.mark.skipif(not R.dbs_available(), reason="reference DBS build not present")
def test_live_reference_full_bands():
    t = np.concatenate([np.zeros(9), np.geomspace(0.01, 30.0, 40), np.full(8, 30.0)])
    ks, ns = 9, len(t) - 9
    ref = R.RefDbs(ks, ns, 8, 9, t)
    g = O.DbsGrid(ks, ns, 8, 9, t)
    for f in range(4):
        for k in (0, 4):
            a, b = g.mrk(f, k), ref.mrk(f, k)
            m = b != 0
            assert np.array_equal(a != 0, m)
            assert np.max(np.abs(a[m] - b[m]) / b[m]) < 1e-14
