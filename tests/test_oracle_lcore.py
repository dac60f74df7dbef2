"""Pins of the oracle's one-electron operator with the core (Gen_Lval /
INT_core, src/bsr/BSR_MAT3/l_core.f90) and of its L_data (L_data.f90):

  * ZCB against sympy's 3j symbols;
  * B-spline first and second derivatives against scipy's BSpline, and
    <B_i|B''_j> (mdb2) against exact integrals of the same splines, with the
    Bloch term of INT_core against its closed form -B'_{ns-1}(R);
  * the hydrogenic spectrum of the bare operator (no core): the generalized
    eigenvalues of (L, S) are Z^2/n^2;
  * hl_full, L_int and L_vec against the reference's own Gen_Lval
    (fixture tests/golden/lcore_vectors.npz, and the live build when
    oracle/_ref is present);
  * every L_data structure against a dense restatement of L_data's table
    from hl_full, L_int, L_vec and GET_V.
"""
import os

import numpy as np
import pytest
from scipy.interpolate import BSpline
from scipy.linalg import eigh

from oracle import oracle as O
from oracle import refbind as R

GOLD = os.path.join(os.path.dirname(__file__), "golden", "lcore_vectors.npz")


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _grid(ks=8, ns=40, z=6.0, rmax=15.0, h=0.1):
    from bsr_amd import grid as G
    gg = G.fixed_ns_grid(ks, ns, z, rmax, h)
    return gg, O.Grid(ks, gg.ns, gg.ml, gg.me, gg.h, gg.t)


def test_zcb_is_the_squared_3j_symbol():
    from sympy.physics.wigner import wigner_3j
    for k1 in range(5):
        for k2 in range(8):
            for k3 in range(5):
                want = float(wigner_3j(k1, k2, k3, 0, 0, 0) ** 2)
                assert abs(O.zcb(k1, k2, k3) - want) <= 1e-15 * max(want, 1.0), (k1, k2, k3)


def test_bspline_derivatives_against_scipy():
    gg, g = _grid()
    ks, ns, nv = g.ks, g.ns, g.nv
    d1, d2 = g.bspd()
    gr, _, bsp, _ = g.tables()
    for ith in range(ks):
        for iv in range(nv):
            c = np.zeros(ns)
            c[iv + ith] = 1.0
            b = BSpline(gg.t, c, ks - 1)
            x = gr[iv]
            np.testing.assert_allclose(bsp[iv, :, ith], b(x), atol=1e-13)
            np.testing.assert_allclose(d1[iv, :, ith], b.derivative(1)(x), rtol=1e-11,
                                       atol=1e-11 * np.abs(d1).max())
            np.testing.assert_allclose(d2[iv, :, ith], b.derivative(2)(x), rtol=1e-10,
                                       atol=1e-11 * np.abs(d2).max())


def _exact_db2(gg, ns, ks):
    """<B_i|B''_j> of the spline basis by a 2ks-point Gauss rule per knot
    interval (exact for the degree 2ks-4 integrand), dense (ns, ns)."""
    x0, w0 = np.polynomial.legendre.leggauss(2 * ks)
    t = gg.t
    eye = np.eye(ns)
    B = [BSpline(t, eye[i], ks - 1) for i in range(ns)]
    B2 = [b.derivative(2) for b in B]
    D = np.zeros((ns, ns))
    for iv in range(ks - 1, ns):
        a, b = t[iv], t[iv + 1]
        if b <= a:
            continue
        x = 0.5 * (b - a) * x0 + 0.5 * (b + a)
        w = 0.5 * (b - a) * w0
        V = np.array([f(x) for f in B])
        V2 = np.array([f(x) for f in B2])
        D += (V * w) @ V2.T
    return D


def test_db2_is_the_exact_integral_with_the_bloch_slot():
    gg, g = _grid()
    ns, ks = g.ns, g.ks
    db2 = g.db2()
    D = _exact_db2(gg, ns, ks)
    scale = np.abs(D).max()
    for i in range(ns):
        for jp in range(ks):
            j = i + jp - (ks - 1)
            if j >= 0:
                assert abs(db2[i, jp] - D[i, j]) <= 1e-11 * scale, (i, jp)
    # slot (1,1): <B_{ns-1}|B''_ns>, and the Bloch term of INT_core
    assert abs(db2[0, 0] - D[ns - 2, ns - 1]) <= 1e-11 * scale
    c = np.zeros(ns)
    c[ns - 2] = 1.0
    dB = BSpline(gg.t, c, ks - 1).derivative(1)(gg.t[ns])
    bloch = db2[0, 0] - db2[ns - 1, ks - 2]
    assert abs(bloch - (-dB)) <= 1e-10 * abs(dB)


def _dense_lower(band):
    ns, ks = band.shape
    A = np.zeros((ns, ns))
    for i in range(ns):
        for jp in range(ks):
            j = i + jp - (ks - 1)
            if j >= 0:
                A[i, j] = A[j, i] = band[i, jp]
    return A


def test_bare_operator_has_the_hydrogenic_spectrum():
    """L = d2/dr2 - l(l+1)/r^2 + 2Z/r without a core: L c = lam S c on the
    splines beyond the first (p(0) = 0) has lam = Z^2/n^2, n > l."""
    from bsr_amd import grid as G
    Z = 2.0
    gg = G.fixed_ns_grid(8, 90, Z, 40.0, 0.05)
    g = O.Grid(8, gg.ns, gg.ml, gg.me, gg.h, gg.t)
    ns = g.ns
    nbf = 3
    rng = np.random.default_rng(1)
    pbs = rng.standard_normal((ns, nbf))
    pbs[0] = 0.0

    class J:
        pass
    j = J()
    j.ns, j.nbf, j.pbs = ns, nbf, pbs
    j.iech = np.zeros(nbf, np.int32)
    j.lbs = np.array([0, 1, 2], np.int32)
    j.ibort = np.ones((nbf, nbf), np.int32)
    j.nch, j.ncp, j.mk, j.eps_ovl, j.eps_c = 1, 0, 0, 1e-8, 0.0
    w = O.Wave(g, j)
    w.gen_lval(Z, 0)
    mlh, hl, _, _ = w.lval()
    _, _, _, sb = g.tables()
    S = _dense_lower(sb)
    for l in range(mlh + 1):
        lam = eigh(hl[1:, 1:, l], S[1:, 1:], eigvals_only=True)[::-1]
        want = np.array([Z * Z / n ** 2 for n in range(l + 1, l + 3)])  # states well inside the box
        np.testing.assert_allclose(lam[:2], want, rtol=1e-7, err_msg=f"l={l}")


def _job(pbs, lbs, iech, nch=2, ncp=2, mk=2):
    class J:
        pass
    j = J()
    j.ns, j.nbf = pbs.shape
    j.pbs = pbs
    j.lbs = np.asarray(lbs, np.int32)
    j.iech = np.asarray(iech, np.int32)
    j.ibort = np.ones((j.nbf, j.nbf), np.int32)
    j.nch, j.ncp, j.mk, j.eps_ovl, j.eps_c = nch, ncp, mk, 1e-8, 0.0
    return j


@pytest.mark.parametrize("core", ["k0", "core"])
def test_gen_lval_matches_reference_fixture(core):
    f = np.load(GOLD)
    g = O.Grid(int(f["ks"]), int(f["ns"]), int(f["ml"]), int(f["me"]), float(f["h"]), f["t"])
    kc = 0 if core == "k0" else int(f["kclosd"])
    np.testing.assert_allclose(g.db2(), f["db2"], rtol=0, atol=1e-13 * np.abs(f["db2"]).max())
    w = O.Wave(g, _job(f["pbs"], f["lbs"], f["iech"]))
    w.gen_lval(float(f["z"]), kc)
    mlh, hl, li, lv = w.lval()
    want = f[f"hl_full_k{kc}"]
    assert mlh + 1 == want.shape[2]
    for l in range(mlh + 1):
        assert rel(hl[:, :, l], want[:, :, l]) < 1e-13, l
    assert rel(li, f[f"lint_k{kc}"]) < 1e-13
    assert rel(lv, f[f"lvec_k{kc}"]) < 1e-13
    if kc:  # the core changes the operator, also outside the band (exchange)
        ns, ks = hl.shape[0], int(f["ks"])
        off = np.abs(np.subtract.outer(np.arange(ns), np.arange(ns))) >= ks
        assert rel(hl, f["hl_full_k0"]) > 1e-5
        assert np.abs(hl[off]).max() > 1e-3 and not f["hl_full_k0"][off].any()


@pytest.mark.skipif(not R.available(), reason="reference build oracle/_ref not present")
def test_gen_lval_matches_live_reference():
    gg, g = _grid(ns=60, z=30.0, rmax=20.0, h=0.05)
    rng = np.random.default_rng(5)
    pbs = rng.standard_normal((g.ns, 8))
    pbs[0] = 0.0
    lbs = [0, 1, 2, 0, 1, 3, 2, 0]
    iech = [0, 0, 0, 0, 0, 0, 0, 2]
    r = R.RefGrid(g.ks, gg.ns, gg.ml, gg.me, gg.h, gg.t)
    w = O.Wave(g, _job(pbs, lbs, iech))
    for kc in (0, 1, 4):
        w.gen_lval(30.0, kc)
        mlh, hl, li, lv = w.lval()
        mr, hr, lr, vr = r.gen_lval(pbs, np.array(lbs), np.array(iech), 30.0, kc, g.db2())
        assert mlh == mr
        for l in range(mlh + 1):
            assert rel(hl[:, :, l], hr[:, :, l]) < 1e-13, (kc, l)
        assert rel(li, lr) < 1e-13 and rel(lv, vr) < 1e-13


def test_gen_lval_invariants():
    gg, g = _grid()
    rng = np.random.default_rng(2)
    pbs = rng.standard_normal((g.ns, 6))
    pbs[0] = 0.0
    lbs = [0, 1, 0, 1, 2, 0]
    w = O.Wave(g, _job(pbs, lbs, [0, 0, 0, 0, 0, 3]))
    w.gen_lval(6.0, 2)
    mlh, hl, li, lv = w.lval()
    ns, ks = g.ns, g.ks
    band = np.abs(np.subtract.outer(np.arange(ns), np.arange(ns))) < ks
    for l in range(mlh + 1):
        H = hl[:, :, l]
        assert np.array_equal(H[band], H.T[band])           # the band is mirrored exactly
        assert rel(H, H.T) < 1e-12                           # the core exchange is symmetric
    for i in range(2, 5):                                    # beyond the core, bound
        np.testing.assert_allclose(lv[:, i], hl[:, :, lbs[i]].T @ pbs[:, i], rtol=1e-12,
                                   atol=1e-12 * np.abs(lv[:, i]).max())
        for j in range(2, 5):
            if lbs[i] == lbs[j]:
                assert abs(li[i, j] - pbs[:, i] @ hl[:, :, lbs[i]] @ pbs[:, j]) <= 1e-11 * abs(li[i, j]) + 1e-12
            else:
                assert li[i, j] == 0.0
    assert not li[:2].any() and not lv[:, :2].any() and not lv[:, 5].any()   # core and channel orbitals


def _dense_ldata(job, g, hl, li, lv, jtype, l, lst, iitar, mats):
    """L_data's table (L_data.f90:6-26) from its definitions: jtype 1 the
    scalar L(a,b) times HB / <i|.> / <i|.><.|j> / <i|j>, jtype 2-5 the vector
    L(.,b) into HV / HW, jtype 6-9 c * hl_full(l) on the block (or the
    target overlap)."""
    P = job.pbs
    ns, nbf = job.ns, job.nbf
    _, _, _, sb = g.tables()
    S = _dense_lower(sb)
    qb = S @ P

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

    for q in range(len(lst.c)):
        K1, K2, K3, K4, c = lst.k1[q], lst.k2[q], lst.k3[q], lst.k4[q], lst.c[q]
        if jtype == 1:
            a, b = divmod(K1, O.IBI)
            C = li[a - 1, b - 1] * c
            ic, jc = -K3, -K4
            if ic > 0 and jc > 0:
                i, j = max(ic, jc), min(ic, jc)
                mats.hbb[i * (i - 1) // 2 + j - 1] += C
            elif ic < 0 and jc > 0:
                i1, i2 = divmod(K3, O.IBO)
                mats.hcb[:, job.iech[i1 - 1] - 1, jc - 1] += C * getv(i1, i2)
            elif ic < 0 and jc < 0:
                i1, i2 = divmod(K3, O.IBO)
                j1, j2 = divmod(K4, O.IBO)
                add_block(job.iech[i1 - 1], job.iech[j1 - 1], np.outer(C * getv(i1, i2), getv(j1, j2)))
            else:
                i1, i2 = divmod(K3, O.IBO)
                ich, jch = job.iech[i1 - 1], job.iech[i2 - 1]
                if iitar > 0 and ich != jch:
                    add_block(ich, jch, C * S)
                mats.htarg[blk(ich, jch)] += C
        elif jtype <= 5:
            v = lv[:, K2 - 1]
            if K4 > 0:
                j1, j2 = divmod(K4, O.IBO)
                add_block(K3, job.iech[j1 - 1], np.outer(v, c * getv(j1, j2)))
            else:
                mats.hcb[:, K3 - 1, -K4 - 1] += c * v
        else:
            ich, jch = divmod(K1, O.IBI)
            if ich == jch or iitar > 1:
                add_block(ich, jch, c * hl[:, :, l])
            else:
                mats.otarg[blk(ich, jch)] += c


@pytest.mark.parametrize("iitar", [0, 1, 2])
@pytest.mark.parametrize("jtype", range(1, 10))
def test_ldata_dense_reconstruction(jtype, iitar):
    from bsr_amd import synth
    w = synth.make_wave(52, 5, 16, 3.0, 8.0, nch=3, ncp=3, mk=1, nb=4, ntarg=2, fill=0.4, h=0.25, name="tiny")
    g = O.Grid(w.ks, w.ns, w.ml, w.me, w.h, w.t)
    ow = O.Wave(g, w)
    ow.gen_lval(3.0, 1)
    _, hl, li, lv = ow.lval()
    dense = O.WaveMatrices(w.ns, w.nch, w.ncp, w.mk)
    lists = synth.l_lists(w, 7, fill=0.5)
    n = 0
    for (jt, l), lst in sorted(lists.items()):
        if jt != jtype or len(lst) == 0:
            continue
        n += len(lst)
        ow.ldata(jt, l, lst, iitar=iitar)
        _dense_ldata(w, g, hl, li, lv, jt, l, lst, iitar, dense)
    assert n > 0
    got = ow.get()
    for name in ("hcc", "hcb", "hbb", "htarg", "otarg"):
        b = getattr(dense, name)
        if np.abs(b).max() > 0:
            assert rel(getattr(got, name), b) < 1e-12, name
        else:
            assert np.abs(getattr(got, name)).max() == 0.0, name


def test_ldata_errors():
    from bsr_amd import synth
    w = synth.make_wave(52, 5, 16, 3.0, 8.0, nch=3, ncp=3, mk=1, nb=4, ntarg=2, fill=0.4, h=0.25, name="tiny")
    g = O.Grid(w.ks, w.ns, w.ml, w.me, w.h, w.t)
    ow = O.Wave(g, w)
    lst = synth.RecordList.build([1 * O.IBI + 1], [0], [0], [0], [1.0])
    with pytest.raises(O.OracleError):
        ow.ldata(6, 0, lst)           # Gen_Lval has not run
    ow.gen_lval(3.0, 0)
    with pytest.raises(O.OracleError):
        ow.ldata(6, 99, lst)          # l beyond the orbitals
    with pytest.raises(O.OracleError):
        ow.ldata(1, 0, synth.RecordList.build([1 * O.IBI + 1], [0], [-3], [0], [1.0]))  # unknown structure
