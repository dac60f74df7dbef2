"""GPU parity of the HIP path (through the C ABI) against the CPU oracle.

Tolerance: 1e-12 per element relative to the oracle, with a floor of 1e-3
of the block maximum for elements that cancel (coefficients are signed).
The R^k band itself is non-negative and is compared element by element.
"""
import numpy as np
import pytest

from tests.conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _grids():
    from bsr_amd import grid as G
    out = []
    t, ns, ml, me, h = G.mkgrid(0, 8, 2.0, 0.125, 160.0, 160.0)      # the reference's TEST_BS knot.dat
    out.append(("test_bs", 8, ns, ml, me, h, t))
    g = G.fixed_ns_grid(8, 60, 6.0, 20.0, 0.05)                        # 60-spline grid
    out.append(("ns60", 8, g.ns, g.ml, g.me, g.h, g.t))
    t, ns, ml, me, h = G.mkgrid(1, 9, 3.0, 0.25, 1.5, 15.0)            # ks=9, geometric region
    out.append(("ks9", 9, ns, ml, me, h, t))
    return out


@pytest.mark.parametrize("case", _grids(), ids=lambda c: c[0])
def test_band_and_moments(ctx, oracle_lib, case):
    O = oracle_lib
    name, ks, ns, ml, me, h, t = case
    print(f"[band] grid={name} ks={ks} ns={ns} ml={ml} me={me}", flush=True)
    ctx.set_grid(ks, ns, ml, me, h, t)
    og = O.Grid(ks, ns, ml, me, h, t)
    for k in range(0, 5):
        r1, r2, rd = ctx.rk_moments(k)
        o1, o2, od = og.rk_moments(k)
        assert rel_err(r1, o1, 1e-14) < TOL
        assert rel_err(r2, o2, 1e-14) < TOL
        assert rel_err(rd, od, 1e-14) < TOL
        rkb = ctx.mrk_cell(k)
        orb = og.mrk_cell(k)
        assert np.all(rkb >= 0.0)
        assert rel_err(rkb, orb, 1e-14) < TOL, f"k={k}"


def test_convol_all_cases(ctx, oracle_lib):
    O = oracle_lib
    from bsr_amd import grid as G
    g = G.fixed_ns_grid(8, 60, 6.0, 20.0, 0.05)
    ctx.set_grid(g.ks, g.ns, g.ml, g.me, g.h, g.t)
    og = O.Grid(g.ks, g.ns, g.ml, g.me, g.h, g.t)
    rng = np.random.default_rng(7)
    for k in (0, 3):
        rkb = og.mrk_cell(k)
        p1, p2 = rng.standard_normal(g.ns), rng.standard_normal(g.ns)
        for icase in (1, 2, 3, 4):
            typ = "s" if icase <= 2 else "x"
            d = O.density(g.ns, g.ks, p1, p2, typ)
            ref = O.convol(g.ns, g.ks, d, icase, rkb)
            got = ctx.convol(k, d, icase)
            assert rel_err(got, ref) < TOL, (k, icase)


def _wave(ctx, w):
    ctx.set_grid(w.ks, w.ns, w.ml, w.me, w.h, w.t)
    ctx.set_orbitals(w.pbs, w.iech, w.lbs, w.ibort, w.eps_ovl)


def _cmp_mats(got, ref, what, tol=TOL):
    for name in ("hcc", "hcb", "hbb", "acf", "htarg", "otarg"):
        a, b = getattr(got, name), getattr(ref, name)
        if b.size and np.abs(b).max() > 0:
            e = rel_err(a, b)
            assert e < tol, f"{what}.{name}: {e:.3e}"
        else:
            assert np.abs(a).max() == 0.0, f"{what}.{name} should be zero"


@pytest.mark.parametrize("jtype", range(1, 10))
def test_idata_each_structure(ctx, oracle_lib, jtype):
    """Per-call I_data for one list of each structure vs the oracle."""
    O = oracle_lib
    from bsr_amd import synth
    w = synth.make_wave(11, 8, 40, 4.0, 12.0, nch=6, ncp=7, mk=3, nb=4, ntarg=3, fill=0.3, name="small")
    _wave(ctx, w)
    ctx.allocate_matrix(w.nch, w.ncp, w.mk)
    ow = O.Wave(O.Grid(w.ks, w.ns, w.ml, w.me, w.h, w.t), w)
    for k in range(w.mk + 1):
        lst = w.rk_lists.get((jtype, k))
        if lst is None or len(lst) == 0:
            continue
        print(f"[idata] jtype={jtype} k={k} n={len(lst)}", flush=True)
        ctx.i_data(jtype, k, lst)
        ow.idata(jtype, k, lst)
    _cmp_mats(ctx.get_matrix(), ow.get(), f"jtype{jtype}")


def test_call_by_call_wave(ctx, oracle_lib):
    """The SUB1 sequence through the per-call entry points vs the oracle."""
    O = oracle_lib
    from bsr_amd import synth
    w = synth.make_wave(12, 8, 40, 4.0, 12.0, nch=5, ncp=6, mk=2, nb=3, ntarg=2, fill=0.3, name="small")
    _wave(ctx, w)
    ctx.allocate_matrix(w.nch, w.ncp, w.mk)
    for ich in range(1, w.nch + 1):
        ctx.update_hl(ich, ich, 1.0)
    ctx.o_data(w.olist, w.eps_c)
    ctx.symmetrize_diag()
    S = ctx.get_matrix()
    ctx.scale(w.ec)
    for k in range(w.mk + 1):
        for jt in range(1, 10):
            lst = w.rk_lists.get((jt, k))
            if lst is not None and len(lst):
                ctx.i_data(jt, k, lst)
    for ich in range(1, w.nch + 1):
        ctx.update_hl(ich, ich, w.etarg[w.iptar[ich - 1] - 1] - w.ec)
    ctx.symmetrize_diag()
    ctx.acf_finish(0)
    H = ctx.get_matrix()
    oS, oH = O.run_wave(w)
    _cmp_mats(S, oS, "S")
    _cmp_mats(H, oH, "H")


def test_fused_wave_c1(ctx, oracle_lib, tmp_path):
    """Config 1 (ks=8, ns=60, 20 channels, k<=4) fused on the GPU vs the
    oracle: every element of S and H, ACF, target matrices, the bsr_mat
    file, and the lowest 10 generalized eigenvalues (<= 1e-9)."""
    O = oracle_lib
    from bsr_amd import synth
    import scipy.linalg as sl
    w = synth.c1()
    print(f"[c1] nhm={w.nhm} records={w.n_records()}", flush=True)
    ctx.wave_assemble(w)
    S, H = ctx.wave_get(0), ctx.wave_get(1)
    oS, oH = O.run_wave(w)
    for name in ("hcc",):
        assert rel_err(getattr(S, name), getattr(oS, name)) < TOL
        assert rel_err(getattr(H, name), getattr(oH, name)) < TOL
    assert rel_err(H.acf, oH.acf) < TOL
    assert rel_err(H.htarg, oH.htarg) < TOL
    assert rel_err(S.otarg, oS.otarg) < TOL
    # the file
    p1, p2 = tmp_path / "gpu.mat", tmp_path / "oracle.mat"
    ctx.wave_write(p1, w)
    O.write_bsr_mat(p2, w, oS, oH)
    g, r = O.read_bsr_mat(p1), O.read_bsr_mat(p2)
    assert (g["ns"], g["nch"], g["ncp"], g["mk"]) == (r["ns"], r["nch"], r["ncp"], r["mk"])
    for part, ref in (("S", oS), ("H", oH)):
        scale = np.abs(ref.hcc).max()   # floor relative to the whole matrix
        assert g[part]["idiag"] == r[part]["idiag"]
        assert sorted(g[part]["blocks"]) == sorted(r[part]["blocks"])
        blocks = list(r[part]["blocks"].items()) + [((i, i), b) for i, b in enumerate(r[part]["diag"])]
        gb = dict(g[part]["blocks"])
        gb.update({(i, i): b for i, b in enumerate(g[part]["diag"])})
        worst = 0.0
        for key, b in blocks:
            den = np.maximum(np.abs(b), 1e-3 * scale)
            worst = max(worst, float(np.max(np.abs(gb[key] - b) / den)))
        print(f"[c1] file {part}: max element rel err {worst:.2e}", flush=True)
        assert worst < TOL
    assert rel_err(g["acf"], r["acf"]) < TOL
    # eigenvalues of H c = E S c, lowest 10
    e_g = sl.eigh(H.dense(), S.dense(), eigvals_only=True, subset_by_index=[0, 9])
    e_o = sl.eigh(O.dense_from_wave(oH), O.dense_from_wave(oS), eigvals_only=True, subset_by_index=[0, 9])
    print("[c1] eig", e_o[:3], np.abs(e_g - e_o).max(), flush=True)
    assert np.abs(e_g - e_o).max() <= 1e-9
    # the same eigenvalues from the device solver (wave_eigen, rocSOLVER)
    e_d = ctx.wave_eigen(10)
    print("[c1] device eig err", np.abs(e_d - e_o).max(), flush=True)
    assert np.abs(e_d - e_o).max() <= 1e-9
    # and from the file, read as BSR_HD3 reads it
    e_f = ctx.matfile_eigen(p1, 10)
    print("[c1] file eig err", np.abs(e_f - e_o).max(), flush=True)
    assert np.abs(e_f - e_o).max() <= 1e-9
    assert np.abs(e_f - e_d).max() <= 1e-12


def test_fused_wave_with_perturbers(ctx, oracle_lib):
    """Production-shaped wave (ks=8, ns=200, k<=10, perturbers) with few
    channels so the oracle finishes in seconds."""
    O = oracle_lib
    from bsr_amd import synth
    w = synth.make_wave(21, 8, 200, 18.0, 40.0, nch=4, ncp=12, mk=10, nb=5, ntarg=5, fill=0.05, name="c2s")
    print(f"[c2-small] nhm={w.nhm} records={w.n_records()}", flush=True)
    ctx.wave_assemble(w)
    S, H = ctx.wave_get(0), ctx.wave_get(1)
    oS, oH = O.run_wave(w)
    _cmp_mats(S, oS, "S")
    _cmp_mats(H, oH, "H")
    # lowest 10 eigenvalues with the perturber rows and columns in the basis
    import scipy.linalg as sl
    e_o = sl.eigh(O.dense_from_wave(oH), O.dense_from_wave(oS), eigvals_only=True, subset_by_index=[0, 9])
    e_d = ctx.wave_eigen(10)
    print("[c2-small] device eig err", np.abs(e_d - e_o).max(), flush=True)
    assert np.abs(e_d - e_o).max() <= 1e-9


def test_fused_equals_call_by_call_and_is_deterministic(ctx):
    from bsr_amd import synth
    w = synth.make_wave(31, 8, 80, 10.0, 25.0, nch=12, ncp=30, mk=6, nb=5, ntarg=4, fill=0.1, name="mid")
    ctx.wave_assemble(w)
    H1 = ctx.wave_get(1)
    ctx.wave_assemble(w, set_inputs=False)
    H2 = ctx.wave_get(1)
    assert np.array_equal(H1.hcc, H2.hcc) and np.array_equal(H1.hcb, H2.hcb)
    # diagonal blocks are symmetric
    for i in range(1, w.nch + 1):
        b = H1.hcc[:, :, i * (i + 1) // 2 - 1]
        assert np.array_equal(b, b.T)


def test_wave_sequence_on_one_context(ctx, oracle_lib):
    """Waves of different channel counts in a row on one context (as the
    serving loop runs them): S blocks that hold zeros are not rewritten when
    a wave leaves them empty, so each wave must still match the oracle in
    every block, including blocks a previous wave had filled."""
    O = oracle_lib
    from bsr_amd import synth
    skipped = 0
    for seed, nch in ((91, 7), (92, 4), (93, 7), (94, 6)):
        w = synth.make_wave(seed, 8, 40, 6.0, 15.0, nch=nch, ncp=3, mk=3, nb=3, ntarg=3, fill=0.3, name="seq")
        print(f"[seq] seed={seed} nch={nch}", flush=True)
        ctx.wave_assemble(w)
        cnt = ctx.wave_term_counts(0)
        assert cnt["blocks"] == nch * (nch + 1) // 2
        # after the launch exactly the S blocks with terms may hold nonzeros
        assert max(cnt["band_blocks"], cnt["rank1_blocks"]) <= cnt["live_blocks"]
        assert cnt["live_blocks"] <= min(cnt["blocks"], cnt["band_blocks"] + cnt["rank1_blocks"])
        skipped += cnt["blocks"] - cnt["live_blocks"]
        S, H = ctx.wave_get(0), ctx.wave_get(1)
        oS, oH = O.run_wave(w)
        _cmp_mats(S, oS, "S")
        _cmp_mats(H, oH, "H")
    assert skipped > 0  # the sequence exercises empty S blocks


def test_errors_are_reported(ctx):
    from bsr_amd import synth, BsrGpuError
    w = synth.make_wave(41, 8, 30, 4.0, 10.0, nch=3, ncp=2, mk=1, nb=2, ntarg=2, fill=0.5, name="err")
    _wave(ctx, w)
    ctx.allocate_matrix(w.nch, w.ncp, w.mk)
    lst = w.rk_lists[(8, 0)]
    bad = synth.RecordList(lst.k1.copy(), lst.k2, lst.k3, lst.k4, lst.c)
    bad.k1[0] = 9 * synth.IBI + 1
    with pytest.raises(BsrGpuError, match="UPDATE_HX: channel index out of range"):
        ctx.i_data(8, 0, bad)
    with pytest.raises(BsrGpuError, match="unknown case in I_data"):
        ctx.i_data(8, 0, lst, icase=3)
    with pytest.raises(BsrGpuError, match="unknown itype"):
        ctx.i_data(10, 0, lst)
    with pytest.raises(BsrGpuError, match="unknown case"):
        ctx.convol(0, np.zeros((w.ns, w.ns)), 5)
    from bsr_amd import Context
    with Context(0) as fresh:
        with pytest.raises(BsrGpuError, match="no assembled wave"):
            fresh.wave_eigen(10)
        with pytest.raises(BsrGpuError, match="cannot open"):
            fresh.matfile_eigen("/nonexistent/bsr_mat.001", 10)


@pytest.mark.parametrize("full", [False, True], ids=["rk", "l_orth"])
def test_fortran_host_driver(oracle_lib, tmp_path, full):
    """The Fortran host (module bsr_gpu over the C ABI, program bsr_mat_gpu)
    runs a whole wave call by call and writes bsr_mat.nnn with the Fortran
    runtime's own unformatted records; compared with the oracle's file.
    l_orth adds SUB1's one-electron part (Gen_Lval with a closed core, the
    L-integral lists) and BS_ORTH."""
    import subprocess
    from tests.conftest import HAS_GPU
    if not HAS_GPU:
        pytest.skip("no HIP device")
    O = oracle_lib
    from bsr_amd import synth, jobfile
    from bsr_amd.build import DRIVER
    w = synth.make_wave(61, 8, 40, 4.0, 12.0, nch=5, ncp=6, mk=3, nb=3, ntarg=2, fill=0.3, name="fortran")
    extra = {}
    if full:
        w.lbs[:] = 1  # orthogonality conditions for the channels
        llists = synth.l_lists(w, 8, fill=0.3)
        for lst in llists.values():  # coefficients below eps_c, which L_data skips (L_data.f90:51,101,130)
            lst.c[::7] *= 1e-12
        extra = dict(l_lists=llists, z=4.0, kclosd=1, iitar=1)
    job, out, ref = tmp_path / "wave.bin", tmp_path / "bsr_mat.001", tmp_path / "oracle.001"
    jobfile.write(job, w, orth=full, **extra)
    r = subprocess.run([DRIVER, str(job), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    oS, oH = O.run_wave(w, orth=full, **extra)
    O.write_bsr_mat(ref, w, oS, oH)
    g, o = O.read_bsr_mat(out), O.read_bsr_mat(ref)
    assert (g["ns"], g["nch"], g["ncp"], g["mk"], g["ns2"]) == (o["ns"], o["nch"], o["ncp"], o["mk"], o["ns2"])
    for part, M in (("S", oS), ("H", oH)):
        scale = np.abs(M.hcc).max()
        assert sorted(g[part]["blocks"]) == sorted(o[part]["blocks"])
        assert sorted(g[part]["vec"]) == sorted(o[part]["vec"])
        assert sorted(g[part]["scal"]) == sorted(o[part]["scal"])
        assert g[part]["idiag"] == o[part]["idiag"]
        for key, b in list(o[part]["blocks"].items()) + list(enumerate(o[part]["diag"])):
            gb = g[part]["blocks"][key] if isinstance(key, tuple) else g[part]["diag"][key]
            assert np.max(np.abs(gb - b) / np.maximum(np.abs(b), 1e-3 * scale)) < TOL
        for key, v in o[part]["vec"].items():
            assert rel_err(g[part]["vec"][key], v) < TOL
    for name in ("acf", "htarg", "otarg", "etarg"):
        assert rel_err(g[name], o[name]) < TOL
    assert g["ec"] == o["ec"]


@pytest.mark.parametrize("mode,nslices", [(1, 6), (1, 7), (2, 6), (2, 7)])
def test_split_precision_wave(ctx, oracle_lib, mode, nslices):
    """Exchange contraction as Ozaki slices on bf16 (mode 1) or int8 (mode 2)
    MFMA: every element of H within 1e-10 (floor 1e-3 of max|H|) of the FP64
    oracle, S unaffected, and the lowest 10 eigenvalues of H x = E S x
    within 1e-9 of the oracle's (LAPACK)."""
    O = oracle_lib
    from bsr_amd import synth
    w = synth.make_wave(71, 8, 120, 18.0, 40.0, nch=6, ncp=10, mk=6, nb=5, ntarg=5, fill=0.1, name="split")
    try:
        ctx.set_precision(mode, nslices)
        ctx.wave_assemble(w)
        S, H = ctx.wave_get(0), ctx.wave_get(1)
        assert ctx.wave_engine() == 10 * mode + nslices
    finally:
        ctx.set_precision(0, 0)
    oS, oH = O.run_wave(w)
    e = rel_err(H.hcc, oH.hcc)
    print(f"[split] mode={mode} S={nslices} max rel err H.hcc {e:.2e}", flush=True)
    assert e <= 1e-10
    assert rel_err(S.hcc, oS.hcc) < TOL
    assert rel_err(H.hcb, oH.hcb) < TOL
    import scipy.linalg as sl
    e_o = sl.eigh(O.dense_from_wave(oH), O.dense_from_wave(oS), eigvals_only=True, subset_by_index=[0, 9])
    de = np.abs(ctx.wave_eigen(10) - e_o).max()
    print(f"[split] mode={mode} S={nslices} max eigenvalue error {de:.2e}", flush=True)
    assert de <= 1e-9


@pytest.mark.parametrize("method", [1, 2])
def test_eigen_methods_vs_lapack(ctx, oracle_lib, method):
    """The lowest 10 eigenvalues of H x = E S x for a wave of nhm = 2420
    (12 channels of ns = 200 and 20 perturbers) by the dense solver (1) and
    by block Lanczos (2), each within 1e-9 of LAPACK on the same matrices;
    Lanczos reports a residual bound below its tolerance, and the automatic
    choice at this size is Lanczos, which returns the same eigenvalues,
    bit for bit, on a second call."""
    O = oracle_lib
    from bsr_amd import synth
    import scipy.linalg as sl
    w = synth.make_wave(91, 8, 200, 18.0, 40.0, nch=12, ncp=20, mk=10, nb=5, ntarg=5, fill=0.05, name="eig")
    ctx.wave_assemble(w)
    S, H = ctx.wave_get(0), ctx.wave_get(1)
    e_ref = sl.eigh(O.dense_from_wave(H), O.dense_from_wave(S), eigvals_only=True, subset_by_index=[0, 9])
    try:
        ctx.set_eigen_method(method)
        e = ctx.wave_eigen(10)
        info = ctx.eigen_info()
    finally:
        ctx.set_eigen_method(0)
    d = np.abs(e - e_ref).max()
    print(f"[eigen] method {method}: {info}, max |dE| vs LAPACK {d:.2e}", flush=True)
    assert info["method"] == method
    assert d <= 1e-9
    if method == 2:
        assert 0 < info["dim"] < w.nhm and info["resid"] <= 1e-11 * max(1.0, np.abs(e).max())
        assert np.array_equal(ctx.wave_eigen(10), e)
        assert ctx.eigen_info()["method"] == 2   # the automatic choice from nhm 2048 on


@pytest.mark.parametrize("engine", [0, 1, 2, 3, 4])
def test_fp64_gemm_engines(ctx, oracle_lib, engine):
    """Each FP64 engine of the exchange contraction (library GEMM, the three
    MFMA kernel tilings, the product over each block's nonzero coefficients
    fused with the term pass) against the oracle."""
    O = oracle_lib
    from bsr_amd import synth
    w = synth.make_wave(81, 8, 90, 18.0, 40.0, nch=7, ncp=5, mk=4, nb=5, ntarg=4, fill=0.15, name="eng")
    try:
        ctx.set_gemm_engine(engine)
        ctx.wave_assemble(w)
        assert ctx.wave_engine() == engine
        H = ctx.wave_get(1)
    finally:
        ctx.set_gemm_engine(-1)
    oS, oH = O.run_wave(w)
    assert rel_err(H.hcc, oH.hcc) < TOL
    assert rel_err(H.hcb, oH.hcb) < TOL


def test_fused_engine_is_the_default_and_deterministic(ctx):
    """With the default choice the FP64 wave uses engine 4 when its bases fit
    the LDS tile; two runs give the same bits, and the per-block result does
    not depend on the block count (a wave with one channel fewer gives the
    same leading blocks bit for bit, the blocks' sums being independent)."""
    from bsr_amd import synth
    w = synth.make_wave(82, 8, 120, 18.0, 40.0, nch=9, ncp=6, mk=6, nb=5, ntarg=4, fill=0.1, name="eng4")
    ctx.wave_assemble(w)
    assert ctx.wave_engine() == 4
    H1 = ctx.wave_get(1)
    ctx.wave_assemble(w, set_inputs=False)
    H2 = ctx.wave_get(1)
    assert np.array_equal(H1.hcc, H2.hcc) and np.array_equal(H1.hcb, H2.hcb)


@pytest.mark.parametrize("mode,nslices", [(0, 0), (2, 6), (2, 7), (1, 6)])
def test_fused_term_pass_equals_separate(ctx, oracle_lib, mode, nslices):
    """A prepared wave runs the exchange product with the band terms in its
    epilogue and adds the rank-1 terms in a pass of their own; a one-shot
    wave starts the product early and adds the band and rank-1 terms in one
    separate pass.  Both add the same terms in the same order, so S and H
    agree bit for bit, in FP64 (the sparse product) and in the split forms;
    and the fused result matches the oracle."""
    O = oracle_lib
    from bsr_amd import synth
    # ns^2 = 8100 is not a multiple of the 64-position tiles, 45 blocks not a
    # multiple of the split kernel's 64 columns; perturbers give rank-1 terms
    w = synth.make_wave(83, 8, 90, 18.0, 40.0, nch=9, ncp=8, mk=5, nb=5, ntarg=4, fill=0.12, name="fuse")
    try:
        ctx.set_precision(mode, nslices)
        ctx.wave_assemble(w)
        S1, H1 = ctx.wave_get(0), ctx.wave_get(1)
        counts = ctx.wave_term_counts(1)
        assert counts["rank1_terms"] > 0 and counts["band_terms"] > 0, counts
        ctx.wave_prepare(w, set_inputs=False)
        ctx.wave_launch()
        ctx.wave_finish()
        S2, H2 = ctx.wave_get(0), ctx.wave_get(1)
        assert ctx.wave_engine() == (4 if mode == 0 else 10 * mode + nslices)
    finally:
        ctx.set_precision(0, 0)
    for a, b in ((S1, S2), (H1, H2)):
        assert np.array_equal(a.hcc, b.hcc) and np.array_equal(a.hcb, b.hcb)
        assert np.array_equal(a.hbb, b.hbb) and np.array_equal(a.acf, b.acf)
    oS, oH = O.run_wave(w)
    e = rel_err(H2.hcc, oH.hcc)
    print(f"[fuse] mode={mode} S={nslices} H.hcc vs oracle {e:.2e}", flush=True)
    assert e <= (TOL if mode == 0 else 1e-10)
    assert rel_err(S2.hcc, oS.hcc) < TOL


@pytest.mark.parametrize("persistent", [True, False])
def test_batch_graph(oracle_lib, persistent):
    """Uneven waves captured into one hipGraph (their exchange products in
    one persistent launch, or one branch per wave): results match the oracle
    and two replays give identical bits."""
    from tests.conftest import HAS_GPU
    if not HAS_GPU:
        pytest.skip("no HIP device")
    O = oracle_lib
    from bsr_amd import synth
    from bsr_amd.batch import Batch
    waves = [synth.make_wave(90 + j, 8, ns, 18.0, 40.0, nch=nch, ncp=4, mk=3, nb=5, ntarg=3, fill=0.1,
                             name=f"b{j}") for j, (ns, nch) in enumerate([(60, 5), (90, 3), (75, 6)])]
    b = Batch(waves)
    try:
        b.prepare()
        b.capture(persistent=persistent)
        per, tot = b.run()
        assert len(per) == 3 and tot > 0
        first = [c.wave_get(1).hcc.copy() for c in b.ctx]
        b.run()
        for c, h in zip(b.ctx, first):
            assert np.array_equal(c.wave_get(1).hcc, h)
        for c, w in zip(b.ctx, waves):
            oS, oH = O.run_wave(w)
            assert rel_err(c.wave_get(1).hcc, oH.hcc) < TOL
            assert rel_err(c.wave_get(0).hcc, oS.hcc) < TOL
    finally:
        b.close()
