"""GPU parity of the channel-block steps of a Dirac partial wave: the
overlap pass, UPDATE_HX, the energy shift, Diag_channels and
transform_matrix (bsrg_dirac_overlap .. bsrg_dirac_transform), each against
the oracle (oracle/dbsr_diag.py, pinned against the reference build by
tests/test_oracle_dbsr_diag.py) on the same inputs:

  * S of the overlap pass: within 1e-13 of the oracle's (the same sums in
    the same order; fppqq from the library's own B-spline tables);
    eps_c and iitar variants
  * UPDATE_HX and the energy shift: the blocks the GPU had before plus the
    oracle's update, within 1e-13
  * Diag_channels on the GPU's own H and S diagonal blocks: iprm exactly,
    the solution counts exactly, energies within 1e-10 of the largest |E|
    of the channel, vectors within 1e-7 of their largest component (two
    LAPACK-type solvers; nearby eigenvalues make the vectors differ more
    than the energies), and for every GPU vector the generalized residual
    (1e-12 of its scale), the S normalisation (1e-10), the sign rule, zeros
    at the deleted B-splines, the untouched S columns bit for bit
  * transform of S and of H with the GPU's solutions: every element within
    4 ms eps |C_i|^T |H_ij| |C_j| of the oracle's (the rounding bound of
    two chained products of length ms), skipped blocks and vectors
    unchanged, the overlaps matrix within 1e-10.
Waves: small (ns=24), mid (ns=60, perturbers) and the C5 shape reduced to
12 channels (ns=150 per component, ms=300).
"""
import numpy as np
import pytest

from tests.conftest import HAS_GPU, rel_err

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


def _say(msg):
    print(msg, flush=True)


def _wave(name):
    from bsr_amd import synth_dirac as SD
    if name == "small":
        return SD.small()
    if name == "mid":
        return SD.make_dirac_wave(3, 9, 60, 8, 9, 54.0, 30.0, 4, 8, 2, 3, 0.1, name="dirac_mid")
    return SD.c5(nch=12)


def _fpq(w):
    from oracle import dbsr_diag as D
    from oracle import oracle as O
    return D.fppqq(O.DbsGrid(w.ks, w.ns, w.ksp, w.ksq, w.t).tables(), w.ks, w.ns, w.ksp, w.ksq)


@pytest.mark.parametrize("name,eps_c,iitar", [("small", 1e-10, 0), ("small", 1e-3, 2), ("mid", 1e-10, 0),
                                              ("c5", 1e-10, 2)])
def test_overlap_pass(ctx, name, eps_c, iitar):
    from oracle import dbsr_diag as D
    w = _wave(name)
    _say(f"[overlap] {name} eps_c={eps_c} iitar={iitar} ns={w.ns} nch={w.nch} npert={w.npert}")
    ctx.dirac_wave(w)
    ctx.dirac_overlap_wave(w, eps_c=eps_c, iitar=iitar)
    S = ctx.dirac_overlap_get()
    ov = D.Overlap(w, _fpq(w))
    for L in w.olists:
        ov.odata(L.itype, L.k1, L.k2, L.c, eps_c, iitar)
    e = rel_err(S["hch"], ov.hch)
    _say(f"[overlap] {name}: hch {e:.2e} (bit-equal: {np.array_equal(S['hch'], ov.hch)})")
    assert e < 1e-13
    if w.npert:
        assert rel_err(S["hcp"], ov.hcp) < 1e-13
        assert np.array_equal(S["hp"], ov.hp)
    assert np.array_equal(S["otarg"], ov.otarg)


def _prepare(ctx, w):
    """H (R^k) and S on the device; returns H and S before the L(ij) and
    energy updates."""
    ctx.dirac_wave(w)
    ctx.dirac_overlap_wave(w)
    return ctx.dirac_get(), ctx.dirac_overlap_get()


@pytest.mark.parametrize("name", ["small", "mid", "c5"])
def test_update_hx_and_energy_shift(ctx, name):
    from oracle import dbsr_diag as D
    w = _wave(name)
    H0, S = _prepare(ctx, w)
    ctx.dirac_update_hx(1, w.hd, *w.lij)
    H1 = ctx.dirac_get()
    ref = np.array(H0["hch"], copy=True)
    for a, b, m, c in zip(*w.lij):
        ref[:, :, D.dblk(w.nch, a, b)] = ref[:, :, D.dblk(w.nch, a, b)] + c * w.hd[:, :, m - 1]
    assert rel_err(H1["hch"], ref) < 1e-13
    # off-diagonal records: c D^T for ich < jch, c D for ich > jch
    ich, jch = np.array([2, 1], np.int32), np.array([1, 2], np.int32)
    ctx.dirac_update_hx(1, w.hd, ich, jch, np.array([1, 1], np.int32), np.array([0.5, -0.25]))
    H2 = ctx.dirac_get()
    k = D.dblk(w.nch, 2, 1)
    ref[:, :, k] = (ref[:, :, k] + 0.5 * w.hd[:, :, 0]) + (-0.25) * w.hd[:, :, 0].T
    assert rel_err(H2["hch"], ref) < 1e-13
    ctx.dirac_energy_shift(w.ecore, w.etarg)
    H3 = ctx.dirac_get()
    ref[:, :, : w.nch] = D.energy_shift(ref[:, :, : w.nch], S["hch"][:, :, : w.nch], _fpq(w), w.iptar, w.ecore,
                                        w.etarg)
    e = rel_err(H3["hch"], ref)
    _say(f"[update_hx] {name}: energy-shifted H vs oracle {e:.2e}")
    assert e < 1e-13


def _diag_inputs(ctx, w):
    _prepare(ctx, w)
    ctx.dirac_update_hx(1, w.hd, *w.lij)
    ctx.dirac_energy_shift(w.ecore, w.etarg)
    return ctx.dirac_get(), ctx.dirac_overlap_get()


WINDOWS = {"default": (-2 * 137.035999084 ** 2, 1e10, 1e-3), "narrow": (-1e5, 2e5, 3e3)}


@pytest.mark.parametrize("name,window", [("small", "default"), ("mid", "narrow"), ("c5", "narrow"),
                                         ("c5", "default")])
def test_diag_channels(ctx, name, window):
    from oracle import dbsr_diag as D
    w = _wave(name)
    H, S = _diag_inputs(ctx, w)
    edmin, edmax, egap = WINDOWS[window]
    ipsol = ctx.dirac_diag_channels(w.lch, w.kch, edmin=edmin, edmax=edmax, egap=egap)
    diag, ev, iprm = ctx.dirac_solutions()
    t = ctx.dirac_diag_timings()
    _say(f"[diag] {name} {window}: ipsol {ipsol.tolist()[:8]}.. {t[1]:.2f} ms")
    nv = w.ns - w.ks + 1
    ms = w.ms
    iprm_o, _ = D.zero_cond(w.ns, nv + w.ksp - 1, nv + w.ksq - 1, w.ksp, w.ksq, w.lch, w.kch)
    assert np.array_equal(iprm, iprm_o)
    hd, sd = H["hch"][:, :, : w.nch], S["hch"][:, :, : w.nch]
    dref, ipref = D.diag_channels(hd, sd, iprm_o, w.nch, edmin, edmax, egap)
    assert np.array_equal(ipsol, ipref)
    assert np.array_equal(ev, np.concatenate([diag[:k, ms - 1, i] for i, k in enumerate(ipsol)]))
    worst = 0.0
    for ich in range(w.nch):
        k = ipsol[ich]
        keep = np.nonzero(iprm[ich * ms:(ich + 1) * ms])[0]
        a, b = hd[np.ix_(keep, keep, [ich])][:, :, 0], sd[np.ix_(keep, keep, [ich])][:, :, 0]
        E, Er = diag[:k, ms - 1, ich], dref[:k, ms - 1, ich]
        if k:
            assert np.max(np.abs(E - Er)) <= 1e-10 * np.abs(Er).max(), ich
        X, Xr = diag[:, :k, ich], dref[:, :k, ich]
        if k:
            dv = np.max(np.abs(X - Xr), axis=0) / np.abs(Xr).max(axis=0)
            worst = max(worst, float(dv.max()))
            assert dv.max() < 1e-7, ich
        assert not np.delete(X, keep, axis=0).any()
        Xk = X[keep]
        res = a @ Xk - (b @ Xk) * E
        scale = (np.abs(a).max() + np.abs(E) * np.abs(b).max()) * np.abs(Xk).max(axis=0)
        if k:
            assert np.max(np.abs(res).max(axis=0) / scale) < 1e-12, ich
            assert np.max(np.abs(np.einsum("ik,ij,jk->k", Xk, b, Xk) - 1.0)) < 1e-10, ich
            first = np.argmax(np.abs(X), axis=0)
            assert (X[first, np.arange(k)] > 0).all(), ich
        assert np.array_equal(diag[:, k:ms - 1, ich], sd[:, k:ms - 1, ich])
        assert np.array_equal(diag[k:, ms - 1, ich], sd[k:, ms - 1, ich])
    _say(f"[diag] {name} {window}: vectors vs LAPACK dsygv within {worst:.1e} of their largest component")


def _bound_check(got, ref, before, diag, ipsol, nch, ms):
    tol = 4 * ms * EPS
    worst = 0.0
    for ich in range(2, nch + 1):
        for jch in range(1, ich):
            from oracle import dbsr_diag as D
            k = D.dblk(nch, ich, jch)
            if np.max(np.abs(before[:, :, k])) < 1e-20:
                assert np.array_equal(got[:, :, k], before[:, :, k])
                continue
            i, j = ipsol[ich - 1], ipsol[jch - 1]
            bound = np.abs(diag[:, :i, ich - 1]).T @ np.abs(before[:, :, k]) @ np.abs(diag[:, :j, jch - 1])
            d = np.abs(got[:, :, k] - ref[:, :, k])
            assert not d[i:].any() and not d[:, j:].any(), (ich, jch)
            r = d[:i, :j] / np.maximum(bound, 1e-300)
            worst = max(worst, float(r.max()) if r.size else 0.0)
    assert worst <= tol, worst
    return worst


def _vec_check(got, ref, before, diag, ipsol, nch, npert, ms):
    for ich in range(1, nch + 1):
        for ip in range(1, npert + 1):
            q = (ich - 1) * npert + ip - 1
            if np.max(np.abs(before[:, q])) < 1e-20:
                assert np.array_equal(got[:, q], before[:, q])
                continue
            k = ipsol[ich - 1]
            bound = np.abs(diag[:, :k, ich - 1]).T @ np.abs(before[:, q])
            d = np.abs(got[:, q] - ref[:, q])
            assert not d[k:].any() and np.all(d[:k] <= 4 * ms * EPS * bound + 1e-300), (ich, ip)


@pytest.mark.parametrize("name", ["small", "mid", "c5"])
def test_transform(ctx, name, tmp_path):
    from oracle import dbsr_diag as D
    w = _wave(name)
    H, S = _diag_inputs(ctx, w)
    ipsol = ctx.dirac_diag_channels(w.lch, w.kch, edmin=-1e5, edmax=2e5, egap=3e3)
    diag, _, _ = ctx.dirac_solutions()
    ms = w.ms
    ctx.dirac_transform(0)
    St = ctx.dirac_overlap_get()
    ov = ctx.dirac_overlaps()
    h, v, ovr = D.transform(S["hch"], S["hcp"], diag, ipsol, w.nch, w.npert, S["hp"])
    assert np.array_equal(St["hch"][:, :, : w.nch], S["hch"][:, :, : w.nch])
    ws = _bound_check(St["hch"], h, S["hch"], diag, ipsol, w.nch, ms)
    _vec_check(St["hcp"], v, S["hcp"], diag, ipsol, w.nch, w.npert, ms)
    assert np.array_equal(ov == 0, ovr == 0)
    assert np.allclose(ov, ovr, rtol=1e-10, atol=0)
    path = tmp_path / "dbsr_mat.nnn"
    ctx.dirac_record(path, 0)
    nS = len(D.read_records(path))
    ctx.dirac_transform(1)
    Ht = ctx.dirac_get()
    h, v, _ = D.transform(H["hch"], H["hcp"], diag, ipsol, w.nch, w.npert)
    wh = _bound_check(Ht["hch"], h, H["hch"], diag, ipsol, w.nch, ms)
    _vec_check(Ht["hcp"], v, H["hcp"], diag, ipsol, w.nch, w.npert, ms)
    ctx.dirac_record(path, 1, nelc_core=3)
    _check_file(path, nS, w, diag, ipsol, St, Ht, h, v)
    t = ctx.dirac_diag_timings()
    _say(f"[transform] {name}: S {ws / EPS:.1f} eps, H {wh / EPS:.1f} eps of the bound; "
         f"ms: overlap {t[0]:.2f} diag {t[1]:.2f} S {t[2]:.2f} H {t[3]:.2f}")


def _check_file(path, nS, w, diag, ipsol, St, Ht, h_ora, v_ora):
    """The dbsr_mat.nnn records against the restatements (oracle/dbsr_diag.py,
    Record_matrix pinned against the reference's own by
    tests/test_oracle_dbsr_diag.py): byte for byte on the GPU's own
    transformed matrices; against the oracle's transform the same records
    (same skipped blocks), values within the transform's rounding bound."""
    from oracle import dbsr_diag as D
    got = D.read_records(path)
    nv = w.ns - w.ks + 1
    head = D.solution_records(w.ns, w.nch, w.npert, nv + w.ksp - 1, nv + w.ksq - 1, diag, ipsol)
    recS = D.record_matrix(St["hch"], St["hcp"], St["hp"], ipsol, w.nch, w.npert)
    recH = D.record_matrix(Ht["hch"], Ht["hcp"], Ht["hp"], ipsol, w.nch, w.npert)
    tail = D.acf_records(Ht["acf"].ravel(order="F"), w.nch, w.mk, 3)
    want = head + recS + recH + tail
    assert nS == len(head) + len(recS)
    assert len(got) == len(want)
    for i, (a, b) in enumerate(zip(got, want)):
        assert a == b.tobytes(), i
    ora = D.record_matrix(h_ora, v_ora, Ht["hp"], ipsol, w.nch, w.npert)
    assert len(ora) == len(recH)
    for a, b in zip(recH, ora):
        assert a.dtype == b.dtype and a.shape == b.shape
        if a.dtype == np.int32:
            assert np.array_equal(a, b)
        else:
            # the elements themselves are held to the rounding bound above
            assert np.allclose(a, b, rtol=0, atol=1e-10 * float(np.abs(b).max()))
    _say(f"[record] {w.name}: {len(got)} records, {len(recS) // 2} S and {len(recH) // 2} H pieces")
