"""Pins of the oracle of the channel-block steps of DBSR_MAT3's SUB1
(oracle/dbsr_diag.py; no GPU).

  * against the reference build (oracle/_ref/libdbs_ref.so, the reference's
    own DBS_gauss, O_data, updates, target_subs, zero_cond and
    transform_matrix compiled from its sources), when present:
      - fppqq on the diagonal blocks plus the overlap lists, bit for bit,
        with the default eps_c and with eps_c = 1e-3 and iitar = 2
      - Zero_cond over the l/kappa defaults and the explicit
        ilzero/jlzero/ibzero/jbzero forms, exactly
      - transform_matrix of S and of H and its overlaps matrix
  * Diag_channels, whose DSYGV the build leaves out (the image has no
    LAPACK for Fortran): the reduced problems go to LAPACK's dsygv itself
    (scipy's driver "gv", the routine diag.f90 calls), and the result is
    checked against what fixes it -- the generalized residual and S
    normalisation of every kept vector, zeros at the deleted B-splines, the
    sign rule, the energy window by brute force, the untouched columns of
    the S diagonal block
  * the energy shift against its definition written out per element.
"""
import numpy as np
import pytest

from oracle import dbsr_diag as D
from oracle import oracle as O
from oracle import refbind as R

C2 = 137.035999084 ** 2

needs_ref = pytest.mark.skipif(not R.dbs_available(), reason="reference build oracle/_ref/libdbs_ref.so absent")


def _waves():
    from bsr_amd import synth_dirac as SD
    return {"small": SD.small(), "mid": SD.make_dirac_wave(3, 9, 60, 8, 9, 54.0, 30.0, 4, 8, 2, 3, 0.1,
                                                           name="dirac_mid")}


@pytest.fixture(scope="module")
def waves():
    return _waves()


def _fpq(w, tables=None):
    if tables is None:
        tables = O.DbsGrid(w.ks, w.ns, w.ksp, w.ksq, w.t).tables()
    return D.fppqq(tables, w.ks, w.ns, w.ksp, w.ksq)


def _overlap(w, fpq, eps_c=1e-10, iitar=0):
    ov = D.Overlap(w, fpq)
    for L in w.olists:
        ov.odata(L.itype, L.k1, L.k2, L.c, eps_c, iitar)
    return ov


def _h_diag(w, H, S, fpq):
    """H diagonal blocks: R^k + L(ij) one-electron records + energy shift."""
    hd = np.array(H["hch"][:, :, : w.nch], order="F", copy=True)
    for a, b, m, c in zip(*w.lij):
        assert a == b
        hd[:, :, a - 1] += c * w.hd[:, :, m - 1]
    return D.energy_shift(hd, S["hch"][:, :, : w.nch], fpq, w.iptar, w.ecore, w.etarg)


@needs_ref
@pytest.mark.parametrize("name", ["small", "mid"])
@pytest.mark.parametrize("eps_c,iitar", [(1e-10, 0), (1e-10, 1), (1e-3, 2)])
def test_overlap_pass_vs_reference(waves, name, eps_c, iitar):
    w = waves[name]
    ref = R.dirac_overlap(w, eps_c=eps_c, iitar=iitar)
    for tables in (None, R.RefDbs(w.ks, w.ns, w.ksp, w.ksq, w.t).tables()):
        ov = _overlap(w, _fpq(w, tables), eps_c, iitar)
        assert np.array_equal(ov.hch, ref["hch"]), name
        assert np.array_equal(ov.hcp, ref["hcp"]), name
        assert np.array_equal(ov.hp, ref["hp"]), name
        assert np.array_equal(ov.otarg, ref["otarg"]), name
    # the records the rules act on are there: eps_c drops some, iitar adds fppqq off the diagonal
    base = _overlap(w, _fpq(w))
    if eps_c > 1e-10:
        assert not np.array_equal(base.hch, ov.hch)


def test_fppqq_is_the_bspline_gram_matrix(waves):
    """fppqq from the reference's own tables against ZINTYm's integrand
    integrated exactly: the Gram matrix of the P and Q B-splines (a scipy
    BSpline quadrature of high order), an independent restatement."""
    from scipy.interpolate import BSpline
    w = waves["small"]
    fpq = _fpq(w)
    ns = w.ns
    nv = ns - w.ks + 1
    t = np.asarray(w.t)
    for q, kb in ((0, w.ksp), (1, w.ksq)):
        nsb = nv + kb - 1
        tb = np.concatenate([np.full(kb, t[0]), t[w.ks:ns], np.full(kb, t[ns])])
        xg, wg = np.polynomial.legendre.leggauss(2 * kb)
        G = np.zeros((nsb, nsb))
        for iv in range(nv):
            a, b = t[w.ks - 1 + iv], t[w.ks + iv]
            if b <= a:
                continue
            x = 0.5 * (b - a) * xg + 0.5 * (a + b)
            vals = np.array([BSpline.basis_element(tb[i:i + kb + 1], extrapolate=False)(x) for i in range(nsb)])
            vals = np.nan_to_num(vals)
            G += (vals * (0.5 * (b - a) * wg)) @ vals.T
        blk = fpq[q * ns:q * ns + nsb, q * ns:q * ns + nsb]
        assert np.max(np.abs(blk - G)) < 1e-13 * np.abs(G).max(), q
        assert not fpq[q * ns + nsb:(q + 1) * ns].any()


@needs_ref
def test_zero_cond_vs_reference(waves):
    w = waves["mid"]
    R.RefDbs(w.ks, w.ns, w.ksp, w.ksq, w.t)
    r = R.RefDbsr(w.pq, w.ipbs, w.i_ch, w.j_ch, w.v_ch, w.nch, w.npert, w.mk, w.ntarg, w.iptar)
    nv = w.ns - w.ks + 1
    nsp, nsq = nv + w.ksp - 1, nv + w.ksq - 1
    rng = np.random.default_rng(7)
    cases = [(-1, -1, 1, 1), (-1, -1, 0, 0), (2, 3, 1, 1), (0, 0, 2, 1), (-1, 4, 0, 2)]
    for zc in cases:
        for trial in range(3):
            kap = rng.integers(1, 10, w.nch) * rng.choice([-1, 1], w.nch)   # l up to 9: il > ksp-1 too
            lch = np.where(kap > 0, kap, -kap - 1)
            got, nh = D.zero_cond(w.ns, nsp, nsq, w.ksp, w.ksq, lch, kap, *zc)
            ref, nhr = r.zero_cond(lch, kap, *zc)
            assert np.array_equal(got, ref), zc
            assert nh + w.npert == nhr


def _diag_setup(w):
    return R.dirac_assemble(w), R.dirac_overlap(w)


@needs_ref
@pytest.mark.parametrize("name", ["small", "mid"])
def test_diag_channels_pins(waves, name):
    import scipy.linalg as sl
    w = waves[name]
    H, S = _diag_setup(w)
    fpq = _fpq(w)
    hd = _h_diag(w, H, S, fpq)
    sd = S["hch"][:, :, : w.nch]
    nv = w.ns - w.ks + 1
    iprm, _ = D.zero_cond(w.ns, nv + w.ksp - 1, nv + w.ksq - 1, w.ksp, w.ksq, w.lch, w.kch)
    ms = w.ms
    for edmin, edmax, egap in ((-2 * C2, 1e10, 1e-3), (-1e5, 2e5, 3e3)):
        diag, ipsol = D.diag_channels(hd, sd, iprm, w.nch, edmin, edmax, egap)
        for ich in range(w.nch):
            keep = np.nonzero(iprm[ich * ms:(ich + 1) * ms])[0]
            a, b = hd[np.ix_(keep, keep, [ich])][:, :, 0], sd[np.ix_(keep, keep, [ich])][:, :, 0]
            ev = sl.eigh(a, b, lower=True, eigvals_only=True, driver="gv")
            sel = ev[(ev >= edmin) & (ev <= edmax) & (np.abs(ev) >= egap)]
            k = ipsol[ich]
            assert k == len(sel)
            # (values only: another LAPACK path, equal to rounding)
            assert np.allclose(diag[:k, ms - 1, ich], sel, rtol=0, atol=1e-12 * np.abs(ev).max())
            X = diag[:, :k, ich]
            deleted = np.setdiff1d(np.arange(ms), keep)
            assert not X[deleted].any()
            Xk = X[keep]
            res = a @ Xk - (b @ Xk) * sel
            scale = np.abs(a).max() * np.abs(Xk).max(axis=0) + np.abs(sel) * np.abs(b).max() * np.abs(Xk).max(axis=0)
            assert np.max(np.abs(res).max(axis=0) / scale) < 1e-12
            assert np.max(np.abs(np.einsum("ik,ij,jk->k", Xk, b, Xk) - 1.0)) < 1e-10
            first = np.argmax(np.abs(X), axis=0)
            assert (X[first, np.arange(k)] > 0).all()
            # columns past the solutions keep the S diagonal block, except the energies' column
            assert np.array_equal(diag[:, k:ms - 1, ich], sd[:, k:ms - 1, ich])
            assert np.array_equal(diag[k:, ms - 1, ich], sd[k:, ms - 1, ich])
        if egap > 1:
            assert (ipsol < np.array([len(np.nonzero(iprm[i * ms:(i + 1) * ms])[0]) for i in range(w.nch)])).all()


@needs_ref
@pytest.mark.parametrize("name", ["small", "mid"])
def test_transform_vs_reference(waves, name, tmp_path):
    w = waves[name]
    fpq = _fpq(w)
    H, _ = _diag_setup(w)
    S, r = R.dirac_overlap(w, keep=True)
    hd = _h_diag(w, H, S, fpq)
    nv = w.ns - w.ks + 1
    iprm, _ = D.zero_cond(w.ns, nv + w.ksp - 1, nv + w.ksq - 1, w.ksp, w.ksq, w.lch, w.kch)
    diag, ipsol = D.diag_channels(hd, S["hch"][:, :, : w.nch], iprm, w.nch, -1e5, 2e5, 3e3)
    r.transform(0, diag, ipsol)
    St, ovr = r.get(), r.overlaps(w.nch)
    _assert_record_matrix(r, tmp_path / "s.mat", St, ipsol, w)
    h, v, ov = D.transform(S["hch"], S["hcp"], diag, ipsol, w.nch, w.npert, S["hp"])
    assert np.array_equal(h[:, :, : w.nch], St["hch"][:, :, : w.nch])    # diagonal blocks untouched
    _assert_transform_close(h, St["hch"], S["hch"], diag, ipsol, w.nch)
    _assert_vectors_close(v, St["hcp"], S["hcp"], diag, ipsol, w.nch, w.npert)
    assert np.allclose(ov, ovr, rtol=1e-11, atol=0)
    # H: the reference's R^k matrices in its module, transformed there
    RH = R.RefDbsr(w.pq, w.ipbs, w.i_ch, w.j_ch, w.v_ch, w.nch, w.npert, w.mk, w.ntarg, w.iptar)
    for L in w.lists:
        RH.rdata(L.itype, L.kpol, L.k1, L.k2, L.k3, L.k4, L.c)
    H0 = RH.get()
    RH.transform(1, diag, ipsol)
    Ht = RH.get()
    _assert_record_matrix(RH, tmp_path / "h.mat", Ht, ipsol, w)
    h, v, _ = D.transform(H0["hch"], H0["hcp"], diag, ipsol, w.nch, w.npert)
    _assert_transform_close(h, Ht["hch"], H0["hch"], diag, ipsol, w.nch)
    _assert_vectors_close(v, Ht["hcp"], H0["hcp"], diag, ipsol, w.nch, w.npert)


def _assert_record_matrix(r, path, M, ipsol, w):
    """The reference's Record_matrix (record_matrix.f90) over its transformed
    module matrices, byte for byte against the restatement on the same
    matrices; the skipped blocks are the all-zero ones."""
    r.record(path)
    got = D.read_records(path)
    ref = D.record_matrix(M["hch"], M["hcp"], M["hp"], ipsol, w.nch, w.npert)
    assert len(got) == len(ref)
    for a, b in zip(got, ref):
        assert a == b.tobytes()
    nblk = sum(1 for i in range(2, w.nch + 1) for j in range(1, i)
               if np.abs(M["hch"][:, :, D.dblk(w.nch, i, j)]).max() >= 1e-20 and ipsol[i - 1] and ipsol[j - 1])
    assert sum(1 for a in got if len(a) == 8 and 0 < np.frombuffer(a, np.int32)[0] <= w.nch) == nblk


def _assert_vectors_close(got, ref, before, diag, ipsol, nch, npert):
    for ich in range(1, nch + 1):
        for ip in range(1, npert + 1):
            q = (ich - 1) * npert + ip - 1
            if np.max(np.abs(before[:, q])) < 1e-20:
                assert np.array_equal(got[:, q], before[:, q])
                continue
            k = ipsol[ich - 1]
            bound = np.abs(diag[:, :k, ich - 1]).T @ np.abs(before[:, q])
            d = np.abs(got[:, q] - ref[:, q])
            assert not d[k:].any() and np.all(d[:k] <= 1e-13 * bound + 1e-300), (ich, ip)


def _assert_transform_close(got, ref, before, diag, ipsol, nch):
    """Off-diagonal blocks within 1e-13 of the product's magnitude bound
    |C_i|^T |H_ij| |C_j| (two summation orders of the same products)."""
    for ich in range(2, nch + 1):
        for jch in range(1, ich):
            k = D.dblk(nch, ich, jch)
            if np.max(np.abs(before[:, :, k])) < 1e-20:
                assert np.array_equal(got[:, :, k], before[:, :, k])
                continue
            ci, cj = np.abs(diag[:, : ipsol[ich - 1], ich - 1]), np.abs(diag[:, : ipsol[jch - 1], jch - 1])
            bound = ci.T @ np.abs(before[:, :, k]) @ cj
            d = np.abs(got[:, :, k] - ref[:, :, k])
            assert not d[ipsol[ich - 1]:].any() and not d[:, ipsol[jch - 1]:].any()
            assert np.all(d[: ipsol[ich - 1], : ipsol[jch - 1]] <= 1e-13 * bound + 1e-300), (ich, jch)


def test_energy_shift_definition(waves):
    w = waves["small"]
    rng = np.random.default_rng(3)
    ms = w.ms
    h = rng.standard_normal((ms, ms, w.nch))
    s = rng.standard_normal((ms, ms, w.nch))
    f = rng.standard_normal((ms, ms))
    out = D.energy_shift(h, s, f, w.iptar, w.ecore, w.etarg)
    for ich in range(w.nch):
        for x, y in ((0, 0), (3, 5), (ms - 1, 2)):
            e = (h[x, y, ich] + w.ecore * s[x, y, ich]) + (w.etarg[w.iptar[ich] - 1] - w.ecore) * f[x, y]
            assert out[x, y, ich] == e
