"""Pins of the oracle's O_data, of the asymptotic coefficients (ACF) that
I_data accumulates (I_data.f90:192-216, UPDATE_CF) and of SUB1's ACF finish
(sub1.f90:207-219).  CPU only.

Independent tools: B-splines from scipy (BSpline.design_matrix, not the
oracle's de Boor tables), Gauss-Legendre nodes from numpy, dense matrices.
  * O_data: every structure rebuilt densely from its definition
    (O_data.f90:2-76) with <B|p> from a Gram matrix integrated exactly.
  * ACF: CA = sum c <p_a|r^k|p_b> per channel pair, with the radial moments
    integrated on the reference's own rule (ks Gauss points per interval,
    mrm.f90), independently evaluated; for k <= 1 that rule is exact, which
    is checked against a higher-order rule.
  * the physics of the ACF: beyond the support of the bound orbitals the
    direct channel-channel block of H equals
    sum_k CA_k <B_x|r^-(k+1)|B_z>, the multipole tail whose coefficients
    the ACF records (R^k with r1 > r2 separates exactly).  A dropped record,
    a wrong orbital or multipole in CA fails it.
  * the finish: symmetric, entries below eps_acf = 1e-5 cut to zero, the
    k = 0 core term equal to twice the number of electrons in the closed
    shells (a closed nl shell holds one electron per (m_l, m_s) state), and
    F/2 the coefficient of that tail (both channel orders summed).  The
    factor 2 itself is the reference's convention (parity unpinned).
"""
import numpy as np
import pytest
from scipy.interpolate import BSpline

from oracle import oracle as O


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _moment_matrix(t, ks, ns, m, npts=None):
    """M[i,j] = integral of B_i B_j r^m, npts Gauss-Legendre points per knot
    interval (default ks, the reference's rule)."""
    npts = ks if npts is None else npts
    x, w = np.polynomial.legendre.leggauss(npts)
    M = np.zeros((ns, ns))
    for iv in range(ks - 1, ns):
        a, b = t[iv], t[iv + 1]
        if b <= a:
            continue
        r = 0.5 * (b - a) * x + 0.5 * (b + a)
        wr = 0.5 * (b - a) * w * r ** m
        Bv = BSpline.design_matrix(r, t, ks - 1).toarray()[:, :ns]
        M += Bv.T @ (wr[:, None] * Bv)
    return M


def _tiny(seed=51, mk=1, nb=3, **kw):
    from bsr_amd import synth
    return synth.make_wave(seed, 5, 16, 3.0, 8.0, nch=3, ncp=3, mk=mk, nb=nb, ntarg=2, fill=0.4, h=0.25,
                           name="tiny", **kw)


def _getv_factory(job, S):
    P = job.pbs
    qb = S @ P

    def getv(ii, jj):
        """GET_V, get_V.f90:2-27: <B|p_jj> minus projections on the same-l
        bound orbitals channel orbital ii is not exempt from"""
        v = qb[:, jj - 1].copy()
        for i in range(1, job.nbf + 1):
            if i == jj or job.iech[i - 1] != 0 or job.lbs[i - 1] != job.lbs[ii - 1]:
                continue
            if job.ibort[i - 1, ii - 1] != 0:
                continue
            o = P[:, i - 1] @ S @ P[:, jj - 1]
            if abs(o) < job.eps_ovl:
                continue
            v -= o * qb[:, i - 1]
        return v
    return getv


def _blk(i, j):
    a, b = max(i, j), min(i, j)
    return a * (a - 1) // 2 + b - 1


# ---------------------------------------------------------------- O_data

def test_moment_matrix_exact_gram():
    """The independent Gram matrix: ks points per interval integrate
    B_i B_j exactly (degree 2ks-2), so a higher rule agrees."""
    w = _tiny()
    t, ks, ns = w.t, w.ks, w.ns
    assert rel(_moment_matrix(t, ks, ns, 0), _moment_matrix(t, ks, ns, 0, ks + 3)) < 1e-14
    assert rel(_moment_matrix(t, ks, ns, 1), _moment_matrix(t, ks, ns, 1, ks + 3)) < 1e-14
    # and the oracle's band Gram matrix is the same matrix
    g = O.Grid(w.ks, w.ns, w.ml, w.me, w.h, w.t)
    _, _, _, sb = g.tables()
    S = _moment_matrix(t, ks, ns, 0)
    for i in range(ns):
        for jp in range(ks):
            j = i + jp - (ks - 1)
            if j >= 0:
                assert abs(sb[i, jp] - S[i, j]) < 1e-14 * np.abs(S).max()


@pytest.mark.parametrize("seed", [51, 52])
def test_odata_dense_reconstruction(seed):
    w = _tiny(seed)
    ol = w.olist
    c = ol.c.copy()
    c[::5] *= 1e-12          # below eps_c: O_data skips them (O_data.f90:34)
    from bsr_amd.synth import RecordList
    lst = RecordList(ol.k1, ol.k2, ol.k3, ol.k4, c)
    g = O.Grid(w.ks, w.ns, w.ml, w.me, w.h, w.t)
    ow = O.Wave(g, w)
    ow.odata(lst)
    got = ow.get()
    ns, nch, ncp = w.ns, w.nch, w.ncp
    S = _moment_matrix(w.t, w.ks, ns, 0)
    getv = _getv_factory(w, S)
    hcc = np.zeros((ns, ns, nch * (nch + 1) // 2))
    hcb = np.zeros((ns, nch, ncp))
    hbb = np.zeros(ncp * (ncp + 1) // 2)
    otarg = np.zeros(nch * (nch + 1) // 2)
    seen = set()
    for K3, K4, cc in zip(lst.k3, lst.k4, lst.c):
        if abs(cc) < w.eps_c:
            continue
        ic, jc = -K3, -K4
        if ic > 0 and jc > 0:
            seen.add("HB")
            hbb[_blk(ic, jc)] += cc
        elif ic < 0 and jc > 0:
            seen.add("HV")
            i1, i2 = divmod(K3, O.IBO)
            hcb[:, w.iech[i1 - 1] - 1, jc - 1] += cc * getv(i1, i2)
        elif ic < 0 and jc < 0:
            seen.add("HW")
            i1, i2 = divmod(K3, O.IBO)
            j1, j2 = divmod(K4, O.IBO)
            ich, jch = w.iech[i1 - 1], w.iech[j1 - 1]
            X = np.outer(cc * getv(i1, i2), getv(j1, j2))
            hcc[:, :, _blk(ich, jch)] += X if ich >= jch else X.T
        else:
            seen.add("target")
            i1, i2 = divmod(K3, O.IBO)
            otarg[_blk(w.iech[i1 - 1], w.iech[i2 - 1])] += cc
    assert seen == {"HB", "HV", "HW", "target"}
    assert rel(got.hcc, hcc) < 1e-12
    assert rel(got.hcb, hcb) < 1e-12
    assert rel(got.hbb, hbb) < 1e-12
    assert rel(got.otarg, otarg) < 1e-12
    assert not got.htarg.any() and not got.acf.any()


# ---------------------------------------------------------------- ACF

def _quadr(w, m):
    M = _moment_matrix(w.t, w.ks, w.ns, m)
    return w.pbs.T @ M @ w.pbs


@pytest.mark.parametrize("seed", [53, 54])
def test_idata_acf_dense(seed):
    """Raw ACF after the direct lists (jtype 6, 7) of k = 0..3 equals
    sum c <p_K2|r^k|p_K3> per (k, ich, jch) of the records."""
    w = _tiny(seed, mk=3)
    g = O.Grid(w.ks, w.ns, w.ml, w.me, w.h, w.t)
    ow = O.Wave(g, w)
    acf = np.zeros((w.nch, w.nch, w.mk + 1))
    n = 0
    for k in range(w.mk + 1):
        Q = _quadr(w, k)
        for jt in (6, 7):
            lst = w.rk_lists.get((jt, k))
            if lst is None or len(lst) == 0:
                continue
            ow.idata(jt, k, lst)
            for K1, K2, K3, cc in zip(lst.k1, lst.k2, lst.k3, lst.c):
                ich, jch = divmod(K1, O.IBI)
                acf[ich - 1, jch - 1, k] += cc * Q[K2 - 1, K3 - 1]
                n += 1
    assert n > 20
    got = ow.get().acf
    assert rel(got, acf) < 1e-12
    # exchange and the other structures add nothing to the ACF
    ow2 = O.Wave(g, w)
    for k in range(w.mk + 1):
        for jt in (1, 2, 3, 4, 5, 8, 9):
            lst = w.rk_lists.get((jt, k))
            if lst is not None and len(lst):
                ow2.idata(jt, k, lst)
    assert not ow2.get().acf.any()


def _confined_wave(seed, mk):
    """Tiny wave whose bound orbitals vanish beyond spline r0 (so R^k with
    one electron beyond them separates exactly)."""
    from bsr_amd import synth
    w = synth.make_wave(seed, 5, 24, 3.0, 10.0, nch=3, ncp=0, mk=mk, nb=3, ntarg=2, fill=0.4, h=0.25, name="conf")
    r0 = 10
    w.pbs[r0:, :w.nb] = 0.0
    return w, r0


@pytest.mark.parametrize("seed", [55, 56])
def test_acf_is_the_multipole_tail(seed):
    w, r0 = _confined_wave(seed, mk=3)
    g = O.Grid(w.ks, w.ns, w.ml, w.me, w.h, w.t)
    ow = O.Wave(g, w)
    for k in range(w.mk + 1):
        for jt in (6, 7):
            lst = w.rk_lists.get((jt, k))
            if lst is not None and len(lst):
                ow.idata(jt, k, lst)
    got = ow.get()
    raw = np.moveaxis(got.acf, 2, 0)   # [k, i, j]
    Minv = [_moment_matrix(w.t, w.ks, w.ns, -(k + 1)) for k in range(w.mk + 1)]
    lo = r0 + w.ks - 1          # B_x B_z beyond every bound orbital
    far = slice(lo, w.ns)
    checked = 0
    for i in range(1, w.nch + 1):
        for j in range(1, i + 1):
            tail = np.zeros((w.ns - lo, w.ns - lo))
            for k in range(w.mk + 1):
                ca = raw[k, i - 1, j - 1] + (raw[k, j - 1, i - 1] if i != j else 0.0)
                tail += ca * Minv[k][far, far]
            blk = got.hcc[:, :, _blk(i, j)][far, far]
            if np.abs(tail).max() == 0:
                continue
            assert np.max(np.abs(blk - tail)) < 1e-11 * np.abs(tail).max(), (i, j)
            checked += 1
    assert checked >= 3
    # the finished ACF divided by 2 is that tail's coefficient
    fin = O.Wave(g, w)
    for k in range(w.mk + 1):
        for jt in (6, 7):
            lst = w.rk_lists.get((jt, k))
            if lst is not None and len(lst):
                fin.idata(jt, k, lst)
    fin.acf_finish(0)
    F = np.moveaxis(fin.get().acf, 2, 0)
    for k in range(w.mk + 1):
        for i in range(w.nch):
            for j in range(w.nch):
                ca = raw[k, i, j] + raw[k, j, i] if i != j else raw[k, i, i]
                want = 2.0 * ca if abs(2.0 * ca) >= 1e-5 else 0.0
                assert abs(F[k, i, j] - want) <= 1e-14 * max(1.0, abs(want))


def _closed_shell_electrons(l):
    """States (m_l, m_s) of one nl shell."""
    return len([(ml, ms) for ml in range(-l, l + 1) for ms in (-0.5, 0.5)])


@pytest.mark.parametrize("kclosd", [0, 1, 3])
def test_acf_finish_properties(kclosd):
    w = _tiny(57, mk=2)
    g = O.Grid(w.ks, w.ns, w.ml, w.me, w.h, w.t)
    ow = O.Wave(g, w)
    for k in range(w.mk + 1):
        for jt in (6, 7):
            lst = w.rk_lists.get((jt, k))
            if lst is not None and len(lst):
                ow.idata(jt, k, lst)
    raw = np.moveaxis(ow.get().acf, 2, 0).copy()
    kcore = O.core_acf(w.lbs, kclosd)
    assert kcore == 2 * sum(_closed_shell_electrons(int(w.lbs[i])) for i in range(kclosd))
    ow.acf_finish(kcore)
    F = np.moveaxis(ow.get().acf, 2, 0)
    for k in range(w.mk + 1):
        assert np.array_equal(F[k], F[k].T)
        core = kcore if k == 0 else 0
        off = F[k] - core * np.eye(w.nch)
        assert np.all((off == 0) | (np.abs(off) >= 1e-5))
        # zero exactly where twice the summed coefficient is below the cut
        tot = raw[k] + raw[k].T
        np.fill_diagonal(tot, np.diag(raw[k]))
        assert np.array_equal(off == 0, np.abs(2 * tot) < 1e-5)
