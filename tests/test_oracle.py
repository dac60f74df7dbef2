"""Pins of the CPU oracle (runs without a GPU).

  * known answers of the reference's own BS test (TEST_BS/int_out, knot.dat)
  * golden vectors written by the reference's Fortran routines
    (tests/golden/ref_vectors.npz, see make_golden.py)
  * mathematical invariants (symmetry, positivity, scaling law, icase 3 = 4)
  * a dense reconstruction of every I_data structure from the 4-index
    tensor K(x,z,y,w) = R^k(B_x B_z; B_y B_w), independent of the band
    loops of convol/UPDATE_*
  * the live reference build, when oracle/_ref is present
"""
import json
import os
from math import factorial

import numpy as np
import pytest

from oracle import oracle as O
from oracle import refbind as R

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


# ---------------------------------------------------------------- known answers

def _hydrogenic(n, l, Z, r):
    from scipy.special import eval_genlaguerre
    N = np.sqrt((2 * Z / n) ** 3 * factorial(n - l - 1) / (2 * n * factorial(n + l)))
    return r * N * np.exp(-Z * r / n) * (2 * Z * r / n) ** l * eval_genlaguerre(n - l - 1, 2 * l + 1, 2 * Z * r / n)


def _project(g, n, l, Z):
    """Galerkin projection with p(1..l+1)=0, as the reference's test builds
    its orbitals (bhwf)."""
    gr, grw, bsp, sb = g.tables()
    ns, ks, nv = g.ns, g.ks, g.nv
    yr = _hydrogenic(n, l, Z, gr) * grw
    c = np.zeros(ns)
    for iv in range(nv):
        for ith in range(ks):
            c[iv + ith] += np.sum(bsp[iv, :, ith] * yr[iv, :])
    S = np.zeros((ns, ns))
    for i in range(ns):
        for jp in range(ks):
            j = i + jp - (ks - 1)
            if j >= 0:
                S[i, j] = S[j, i] = sb[i, jp]
    f = l + 1
    x = np.zeros(ns)
    x[f:] = np.linalg.solve(S[f:, f:], c[f:])
    return x


def test_known_answers_int_out():
    """R^k of hydrogenic orbitals (1s with Z=2, the others Z=1) on the
    reference's TEST_BS grid, against the RK_cell column of its int_out."""
    gold = json.load(open(os.path.join(GOLD, "bs_int_out.json")))
    kn = gold["knot"]
    t, ns, ml, me, h = O.mkgrid0(kn["ks"], kn["z"], kn["h"], kn["hmax"], kn["rmax"])
    assert (ns, ml, me, ns - kn["ks"] + 1) == (kn["ns"], kn["ml"], kn["me"], kn["nv"])
    assert np.allclose(np.round(t, 5), kn["t5"], atol=6e-6)
    g = O.Grid(kn["ks"], ns, ml, me, h, t)
    orbs = {}
    for e in gold["int_out"]:
        for name in e["orbs"]:
            if name not in orbs:
                n, l = int(name[0]), "spdf".index(name[1])
                orbs[name] = _project(g, n, l, 2.0 if name == "1s" else 1.0)
    for e in gold["int_out"]:
        a, b, c, d = (orbs[x] for x in e["orbs"])
        rkb = g.mrk_cell(e["k"])
        A = O.density(ns, 8, a, c, "s")
        B = O.density(ns, 8, b, d, "s")
        val = O.sum_amb(ns, 8, O.convol(ns, 8, B, 2, rkb), A, "s")
        assert abs(val - e["rk_cell"]) <= 5e-15 * abs(e["rk_cell"]) + 1e-17, e


def test_closed_form_f0_1s():
    """F0(1s,1s) = 5Z/8 for a hydrogenic 1s orbital (closed form)."""
    t, ns, ml, me, h = O.mkgrid0(8, 2.0, 0.125, 160.0, 160.0)
    g = O.Grid(8, ns, ml, me, h, t)
    for Z in (1.0, 2.0, 3.0):
        p = _project(g, 1, 0, Z)
        d = O.density(ns, 8, p, p, "s")
        val = O.sum_amb(ns, 8, O.convol(ns, 8, d, 1, g.mrk_cell(0)), d, "s")
        assert abs(val - 5 * Z / 8) < 1e-11


# ---------------------------------------------------------------- golden vectors

@pytest.fixture(scope="module")
def gv():
    return np.load(os.path.join(GOLD, "ref_vectors.npz"))


def test_gauss_vs_reference_table(gv):
    """Newton-refined Gauss-Legendre vs the reference's 15-digit table."""
    for k in range(1, 16):
        x, w = O.gauss(k)
        assert np.max(np.abs(x - gv[f"gauss_x_{k}"])) < 2e-15
        assert np.max(np.abs(w - gv[f"gauss_w_{k}"])) < 5e-15
        # exactness on polynomials of degree <= 2k-1
        for deg in range(2 * k):
            assert abs(np.sum(w * x ** deg) - 1.0 / (deg + 1)) < 1e-14


@pytest.mark.parametrize("name", ["g0", "g1", "g2", "g3"])
def test_grids_vs_reference(gv, name):
    kind, ks, z, h, hmax, rmax = gv[f"{name}_args"]
    f = O.mkgrid0 if kind == 0 else O.mkgrid1
    t, ns, ml, me, hh = f(int(ks), z, h, hmax, rmax)
    assert [ns, ml, me] == list(gv[f"{name}_ints"])
    assert np.array_equal(t, gv[f"{name}_t"])
    assert hh == gv[f"{name}_h"][0]


@pytest.mark.parametrize("name", ["g2", "g3"])
def test_integrals_vs_reference(gv, name):
    ks = int(gv[f"{name}_args"][1])
    t = gv[f"{name}_t"]
    ns, ml, me = (int(v) for v in gv[f"{name}_ints"])
    g = O.Grid(ks, ns, ml, me, float(gv[f"{name}_h"][0]), t)
    gr, grw, bsp, sb = g.tables()
    assert rel(gr, gv[f"{name}_gr"]) < 1e-15
    assert rel(grw, gv[f"{name}_grw"]) < 5e-15
    assert rel(bsp, gv[f"{name}_bsp"]) < 5e-15
    assert rel(sb, gv[f"{name}_sb"]) < 5e-15
    p = gv[f"{name}_p"]
    for k in range(4):
        r1, r2, rd = g.rk_moments(k)
        assert rel(r1, gv[f"{name}_rkd1_{k}"]) < 5e-14
        assert rel(r2, gv[f"{name}_rkd2_{k}"]) < 5e-14
        if f"{name}_rkd_{k}" in gv:
            assert rel(rd, gv[f"{name}_rkd_{k}"]) < 5e-14
        rkb = g.mrk_cell(k)
        ref = gv[f"{name}_rkb_{k}"] if f"{name}_rkb_{k}" in gv else None
        if ref is not None:
            m = ref != 0
            assert np.array_equal(rkb != 0, m)
            assert np.max(np.abs(rkb[m] - ref[m]) / ref[m]) < 2e-14
        else:
            sub = rkb.reshape(-1, order="F")[::7]
            refs = gv[f"{name}_rkbs_{k}"]
            m = refs != 0
            assert np.max(np.abs(sub[m] - refs[m]) / refs[m]) < 2e-14
        for icase in (1, 2, 3, 4):
            typ = "s" if icase <= 2 else "x"
            d = O.density(ns, ks, p[:, 0], p[:, 1], typ)
            assert np.array_equal(d, gv[f"{name}_dens_{typ}"])
            assert rel(O.convol(ns, ks, d, icase, rkb), gv[f"{name}_conv_{k}_{icase}"]) < 1e-13
    ds = O.density(ns, ks, p[:, 0], p[:, 1], "s")
    dx = O.density(ns, ks, p[:, 0], p[:, 2], "s")
    assert rel(O.sum_amb(ns, ks, ds, dx, "s"), gv[f"{name}_sumamb_s"][0]) < 1e-14
    assert rel(O.bav(ns, ks, ds, p[:, 2]), gv[f"{name}_bav_s"]) < 1e-14
    assert rel(O.bxv(ks, ns, sb, p[:, 0]), gv[f"{name}_bxv"]) < 1e-14
    assert rel(O.bvmv(ns, ks, sb, p[:, 0], p[:, 1]), gv[f"{name}_bvmv"][0]) < 1e-14
    q = np.array([[g.quadr(p[:, i], p[:, j], m) for m in (0, 1, 2, 3, -1, -2, -3)]
                  for i in range(3) for j in range(3)])
    assert rel(q, gv[f"{name}_quadr"]) < 1e-13


@pytest.mark.parametrize("name", ["g2", "g3"])
def test_accumulators_vs_reference(gv, name):
    """UPDATE_HX/HL/HW/HV/HB/CF of mod_intmat.f90 on a fixed call sequence."""
    from types import SimpleNamespace
    ks = int(gv[f"{name}_args"][1])
    t = gv[f"{name}_t"]
    ns, ml, me = (int(v) for v in gv[f"{name}_ints"])
    g = O.Grid(ks, ns, ml, me, float(gv[f"{name}_h"][0]), t)
    job = SimpleNamespace(nbf=1, pbs=np.zeros((ns, 1)), iech=[0], lbs=[0], ibort=np.zeros((1, 1), np.int32),
                          nch=3, ncp=2, mk=2, eps_ovl=1e-8, eps_c=1e-10, ns=ns)
    w = O.Wave(g, job)
    x = gv[f"{name}_upd_in"]
    dfull = x[: ns * ns].reshape(ns, ns, order="F")
    dband = x[ns * ns: ns * ns + ns * ks].reshape(ns, ks, order="F")
    v = x[ns * ns + ns * ks: ns * ns + ns * ks + ns]
    u = x[ns * ns + ns * ks + ns:]
    w.update_hx(2, 1, dfull, "x")
    w.update_hx(1, 3, dfull, "x")
    w.update_hx(3, 3, dband, "s")
    w.update_hl(2, 2, 0.7)
    w.update_hl(3, 1, -1.3)
    w.update_hw(1, 2, v, u)
    w.update_hw(3, 2, u, v)
    w.update_hv(2, 1, v, 0.5)
    w.update_hb(2, 1, 0.25)
    w.update_cf(1, 3, 1, 2.0)
    m = w.get()
    assert rel(m.hcc, gv[f"{name}_upd_hcc"]) < 1e-15
    assert rel(m.hcb, gv[f"{name}_upd_hcb"]) < 1e-15
    assert rel(m.hbb, gv[f"{name}_upd_hbb"]) < 1e-15
    assert rel(m.acf, gv[f"{name}_upd_acf"]) < 1e-15


# ---------------------------------------------------------------- invariants

def _grid_ns30():
    t, ns, ml, me, h = O.mkgrid1(8, 4.0, 0.25, 1.2, 7.0)
    return O.Grid(8, ns, ml, me, h, t)


def test_band_invariants():
    g = _grid_ns30()
    for k in (0, 2, 5):
        rkb = g.mrk_cell(k)
        assert np.all(rkb >= 0)
        # pair symmetry: rkb(i,j,ip,jp) = rkb(j,i,jp,ip)
        assert rel(rkb, np.transpose(rkb, (1, 0, 3, 2))) < 1e-14
        # nothing outside the band of valid pairs
        ns, ks = g.ns, g.ks
        for ip in range(ks):
            assert np.all(rkb[ns - ip:, :, ip, :] == 0)


def test_scaling_law_matches_direct_quadrature():
    """On a truly geometric region the moment recursion and the direct
    quadrature agree (the recursion is exact in exact arithmetic)."""
    t, ns, ml, me, h = O.mkgrid1(8, 2.0, 0.125, 1.0, 30.0)
    assert me > 20
    ga = O.Grid(8, ns, ml, me, h, t)
    gb = O.Grid(8, ns, ml, 0, h, t)        # me=0: every cell by quadrature
    for k in (0, 1, 4):
        a, b = ga.rk_moments(k), gb.rk_moments(k)
        for x, y in zip(a, b):
            assert rel(x, y) < 1e-12


def test_exchange_cases_coincide():
    g = _grid_ns30()
    rng = np.random.default_rng(3)
    d = rng.standard_normal((g.ns, g.ns))
    rkb = g.mrk_cell(3)
    assert rel(O.convol(g.ns, g.ks, d, 3, rkb), O.convol(g.ns, g.ks, d, 4, rkb)) < 1e-14


# ---------------------------------------------------------------- dense reconstruction

def _K(rkb):
    """Dense K[x,z,y,w] = R^k(B_x B_z; B_y B_w) from the band."""
    ns, _, ks, _ = rkb.shape
    K = np.zeros((ns, ns, ns, ns))
    for x in range(ns):
        for z in range(max(0, x - ks + 1), min(ns, x + ks)):
            for y in range(ns):
                for w in range(max(0, y - ks + 1), min(ns, y + ks)):
                    K[x, z, y, w] = rkb[min(x, z), min(y, w), abs(x - z), abs(y - w)]
    return K


def _dense_idata(job, g, jtype, k, lst, mats):
    """Every I_data structure from its integral definition with the dense
    tensor: jtype 1 R^k(a b; c d) scalars, 2-5 vectors, 6-7 direct band
    blocks, 8-9 exchange blocks (I_data.f90 header table)."""
    K = _K(g.mrk_cell(k))
    P = job.pbs
    ns = job.ns
    _, _, _, sb = g.tables()
    S = np.zeros((ns, ns))
    for i in range(ns):
        for jp in range(g.ks):
            j = i + jp - (g.ks - 1)
            if j >= 0:
                S[i, j] = S[j, i] = sb[i, jp]
    qb = S @ P
    nbf = job.nbf

    def getv(ii, jj):
        v = qb[:, jj - 1].copy()
        for i in range(1, nbf + 1):
            if i == jj or job.iech[i - 1] != 0 or job.lbs[i - 1] != job.lbs[ii - 1]:
                continue
            if job.ibort[i - 1, ii - 1] != 0:
                continue
            o = P[:, i - 1] @ S @ P[:, jj - 1]
            if abs(o) < job.eps_ovl:
                continue
            v -= o * qb[:, i - 1]
        return v

    def blk(i, j):
        a, b = max(i, j), min(i, j)
        return a * (a - 1) // 2 + b - 1

    def add_block(i, j, X):
        mats.hcc[:, :, blk(i, j)] += X if i >= j else X.T

    band = np.abs(np.subtract.outer(np.arange(ns), np.arange(ns))) < g.ks
    for q in range(len(lst.c)):
        K1, K2, K3, K4, c = lst.k1[q], lst.k2[q], lst.k3[q], lst.k4[q], lst.c[q]
        p = lambda o: P[:, o - 1]
        if jtype == 1:
            a, cc = divmod(K1, O.IBI)
            b, d = divmod(K2, O.IBI)
            Sv = np.einsum("xzyw,x,z,y,w->", K, p(a), p(cc), p(b), p(d)) * c
            ic, jc = -K3, -K4
            if ic > 0 and jc > 0:
                i, j = max(ic, jc), min(ic, jc)
                mats.hbb[i * (i - 1) // 2 + j - 1] += Sv
            elif ic < 0 and jc > 0:
                i1, i2 = divmod(K3, O.IBO)
                mats.hcb[:, job.iech[i1 - 1] - 1, jc - 1] += Sv * getv(i1, i2)
            elif ic < 0 and jc < 0:
                i1, i2 = divmod(K3, O.IBO)
                j1, j2 = divmod(K4, O.IBO)
                add_block(job.iech[i1 - 1], job.iech[j1 - 1], np.outer(Sv * getv(i1, i2), getv(j1, j2)))
            else:
                i1, i2 = divmod(K3, O.IBO)
                mats.htarg[blk(job.iech[i1 - 1], job.iech[i2 - 1])] += Sv
        elif jtype <= 5:
            a, b = divmod(K1, O.IBI)
            v = np.einsum("xzyw,z,y,w->x", K, p(K2), p(a), p(b))
            if K4 > 0:
                j1, j2 = divmod(K4, O.IBO)
                add_block(K3, job.iech[j1 - 1], np.outer(v, c * getv(j1, j2)))
            else:
                mats.hcb[:, K3 - 1, -K4 - 1] += c * v
        else:
            ich, jch = divmod(K1, O.IBI)
            if jtype <= 7:
                X = np.einsum("xzyw,y,w->xz", K, p(K2), p(K3)) * band
            else:
                X = np.einsum("xzyw,z,w->xy", K, p(K2), p(K3))
            add_block(ich, jch, c * X)


@pytest.mark.parametrize("jtype", range(1, 10))
def test_idata_dense_reconstruction(jtype):
    from bsr_amd import synth
    w = synth.make_wave(51, 5, 16, 3.0, 8.0, nch=3, ncp=3, mk=1, nb=3, ntarg=2, fill=0.4, h=0.25, name="tiny")
    g = O.Grid(w.ks, w.ns, w.ml, w.me, w.h, w.t)
    ow = O.Wave(g, w)
    dense = O.WaveMatrices(w.ns, w.nch, w.ncp, w.mk)
    n = 0
    for k in range(w.mk + 1):
        lst = w.rk_lists.get((jtype, k))
        if lst is None or len(lst) == 0:
            continue
        n += len(lst)
        ow.idata(jtype, k, lst)
        _dense_idata(w, g, jtype, k, lst, dense)
    assert n > 0
    got = ow.get()
    for name in ("hcc", "hcb", "hbb", "htarg"):
        b = getattr(dense, name)
        if np.abs(b).max() > 0:
            assert rel(getattr(got, name), b) < 1e-12, name
        else:
            assert np.abs(getattr(got, name)).max() == 0.0, name


# ---------------------------------------------------------------- live reference

@pytest.mark.skipif(not R.available(), reason="reference build oracle/_ref not present")
def test_live_reference_c1_grid():
    from bsr_amd import grid as G
    gg = G.fixed_ns_grid(8, 60, 6.0, 20.0, 0.05)
    g = O.Grid(8, gg.ns, gg.ml, gg.me, gg.h, gg.t)
    r = R.RefGrid(8, gg.ns, gg.ml, gg.me, gg.h, gg.t)
    rng = np.random.default_rng(9)
    p = rng.standard_normal((gg.ns, 2))
    for k in range(5):
        a, b = g.mrk_cell(k), r.mrk_cell(k)
        m = b != 0
        assert np.max(np.abs(a[m] - b[m]) / b[m]) < 1e-13
        for icase in (1, 2, 3, 4):
            d = O.density(gg.ns, 8, p[:, 0], p[:, 1], "s" if icase <= 2 else "x")
            assert rel(O.convol(gg.ns, 8, d, icase, a), r.convol(k, d, icase)) < 1e-13
