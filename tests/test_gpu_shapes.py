"""GPU checks at the workloads' real shapes.

  * C2 at full size (ks=8, ns=200, k<=10, 60 channels, 400 perturbers, nhm
    12 400) against a summary of the oracle's own matrices written once by
    tests/golden/make_golden_c2.py (per-block norms of S and H, htarg,
    otarg, a seeded sample of elements, ACF);
  * C3 (batched waves, ns 100..250, k<=8): three waves of the C3 range
    through one hipGraph against the oracle, and the whole 48-wave batch
    replayed from its graph bit for bit equal to its eager run;
  * C4 (served split precision): a stream of C2-shaped jobs served in int8
    Ozaki form, each within 1e-10 of the FP64 form of the same job, and a
    reduced-channel C2-shaped job served in split precision against the
    oracle;
  * C5 (Dirac, ns=150 per component, ks=9 with P order 8, Z=54, k<=12):
    a reduced-channel wave against the oracle in FP64 and its split form
    against FP64.

Tolerances: 1e-12 relative per element for FP64 (floor 1e-3 of the array
maximum), 1e-10 for split precision; norms 1e-12 relative.
"""
import os
import time

import numpy as np
import pytest

from tests.conftest import HAS_GPU, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12
SPLIT_TOL = 1e-10
GOLD = os.path.join(os.path.dirname(__file__), "golden", "c2_full_sample.npz")


def _say(msg):
    print(msg, flush=True)


def _check_summary(M, g, pre, nsamp_check=True):
    """M (host WaveMatrices-like) against the golden summary of prefix pre."""
    hn = np.sqrt(np.einsum("xyb,xyb->b", M.hcc, M.hcc))
    ref = g[f"{pre}_hcc_norm"]
    live = ref > 0
    assert np.array_equal(hn > 0, live), f"{pre}: blocks that hold nonzeros differ"
    assert np.max(np.abs(hn[live] - ref[live]) / ref[live]) < TOL, f"{pre}: block norms"
    fl = 1e-3 * g[f"{pre}_hcc_max"]
    v = M.hcc.ravel(order="F")[g[f"{pre}_hcc_idx"]]
    r = g[f"{pre}_hcc_val"]
    assert np.max(np.abs(v - r) / np.maximum(np.abs(r), fl)) < TOL, f"{pre}: hcc sample"
    for key in ("htarg", "otarg"):
        r = g[f"{pre}_{key}"]
        got = getattr(M, key)
        assert (rel_err(got, r) < TOL) if np.abs(r).max() > 0 else not got.any(), f"{pre}: {key}"
    if f"{pre}_hcb_norm" in g:
        bn = np.sqrt(np.einsum("xcp,xcp->cp", M.hcb, M.hcb))
        ref = g[f"{pre}_hcb_norm"]
        live = ref > 0
        assert np.array_equal(bn > 0, live), f"{pre}: hcb vectors that hold nonzeros differ"
        assert np.max(np.abs(bn[live] - ref[live]) / ref[live]) < TOL, f"{pre}: hcb norms"
        for arr, key in ((M.hcb, "hcb"), (M.hbb, "hbb"), (M.acf, "acf")):
            fl = 1e-3 * g[f"{pre}_{key}_max"]
            v = arr.ravel(order="F")[g[f"{pre}_{key}_idx"]]
            r = g[f"{pre}_{key}_val"]
            if np.abs(r).max() > 0:
                assert np.max(np.abs(v - r) / np.maximum(np.abs(r), fl)) < TOL, f"{pre}: {key} sample"
            else:
                assert not v.any(), f"{pre}: {key} should be zero"


def test_c2_full_size_vs_oracle(ctx):
    from bsr_amd import synth
    g = dict(np.load(GOLD))
    w = synth.c2()
    _say(f"[c2-full] nhm={w.nhm} records={w.n_records()} (golden oracle run {g['seconds']:.0f} s)")
    assert int(g["nhm"]) == w.nhm and int(g["records"]) == w.n_records()
    ctx.wave_assemble(w)
    S = ctx.wave_get(0)
    _check_summary(S, g, "S")
    del S
    H = ctx.wave_get(1)
    _check_summary(H, g, "H")
    _say("[c2-full] S and H match the oracle's block norms and element sample")


def test_c3_range_waves_through_graph(oracle_lib):
    if not HAS_GPU:
        pytest.skip("no HIP device")
    O = oracle_lib
    from bsr_amd import synth
    from bsr_amd.batch import Batch
    waves = [synth.make_wave(300 + ns, 8, ns, 18.0, 40.0, nch=nch, ncp=ncp, mk=8, nb=5, ntarg=5, fill=0.05,
                             name=f"c3-{ns}") for ns, nch, ncp in ((100, 4, 0), (180, 4, 6), (250, 3, 0))]
    b = Batch(waves)
    try:
        b.prepare()
        b.capture()
        per, tot = b.run()
        _say(f"[c3-range] graph of 3 waves {tot:.3f} ms, engines {b.engines()}")
        for c, w in zip(b.ctx, waves):
            oS, oH = O.run_wave(w)
            H, S = c.wave_get(1), c.wave_get(0)
            for name in ("hcc", "hcb", "acf", "htarg"):
                if np.abs(getattr(oH, name)).max() > 0:
                    assert rel_err(getattr(H, name), getattr(oH, name)) < TOL, (w.ns, name)
            assert rel_err(S.hcc, oS.hcc) < TOL, w.ns
            _say(f"[c3-range] ns={w.ns} matches the oracle")
    finally:
        b.close()


def _digest(M):
    import xxhash
    h = xxhash.xxh64()
    for a in (M.hcc, M.hcb, M.hbb, M.acf, M.htarg, M.otarg):
        h.update(np.ascontiguousarray(a).view(np.uint8))
    return h.hexdigest()


def test_c3_batch_graph_equals_eager():
    """The 48 uneven waves of C3 (nch 10..120, ns 100..250, k<=8): the
    graph replay of the batch gives, for every wave, the bits of its eager
    run, with the exchange products of all 48 waves in one persistent launch
    and with one branch per wave."""
    if not HAS_GPU:
        pytest.skip("no HIP device")
    from bsr_amd import synth
    from bsr_amd.batch import Batch
    specs = synth.c3_specs()
    t0 = time.perf_counter()
    waves = [synth.c3_wave(s) for s in specs]
    _say(f"[c3-48] {len(waves)} waves generated in {time.perf_counter() - t0:.1f} s")
    b = Batch(waves)
    try:
        b.prepare()   # one eager pass per wave
        eager = []
        for j, c in enumerate(b.ctx):
            eager.append((_digest(c.wave_get(0)), _digest(c.wave_get(1))))
            if j % 12 == 11:
                _say(f"[c3-48] eager digests {j + 1}/48")
        for persistent in (True, False):
            b.capture(persistent=persistent)
            if persistent:
                nb, bpi = b.batch_info()
                _say(f"[c3-48] persistent launch of {nb} waves, {bpi} blocks per item")
                assert nb == len(waves)
            per, tot = b.run()
            _say(f"[c3-48] graph replay ({'persistent' if persistent else 'per wave'}) {tot:.2f} ms")
            for j, (c, e) in enumerate(zip(b.ctx, eager)):
                assert (_digest(c.wave_get(0)), _digest(c.wave_get(1))) == e, f"wave {j}"
            _say("[c3-48] every wave bit-identical to its eager run")
    finally:
        b.close()


def test_c4_served_split_stream(oracle_lib):
    if not HAS_GPU:
        pytest.skip("no HIP device")
    O = oracle_lib
    from bsr_amd import synth, Context
    from bsr_amd.serve import Server
    stream = synth.c4_stream(njobs=4)
    small = synth.make_wave(21, 8, 200, 18.0, 40.0, nch=4, ncp=12, mk=10, nb=5, ntarg=5, fill=0.05,
                            name="c2s", target_seed=4)
    srv = Server(nslots=3, precision=2, nslices=6)
    try:
        jobs = [srv.submit(w, arrival=0.1 * a, keep=True) for a, w in stream]
        js = srv.submit(small, keep=True)
        for j in jobs + [js]:
            j.event.wait()
            assert j.error is None, j.error
    finally:
        srv.close()
    spans = sorted((j.start, j.done) for j in jobs + [js])
    assert any(b2[0] < a2[1] for a2, b2 in zip(spans, spans[1:]))   # the jobs overlapped
    ref = Context(0)
    try:
        for j in jobs:
            w = j.wave
            ref.wave_assemble(w)
            for which in (0, 1):
                got, fp = j.result[which], ref.wave_get(which)
                e = max(rel_err(got.hcc, fp.hcc), rel_err(got.hcb, fp.hcb))
                assert e < SPLIT_TOL, (w.name, which, e)
            _say(f"[c4] {w.name} nch={w.nch} ncp={w.ncp}: split vs FP64 {e:.2e}")
            j.result = None
    finally:
        ref.close()
    oS, oH = O.run_wave(small)
    S, H = js.result
    e = max(rel_err(H.hcc, oH.hcc), rel_err(H.hcb, oH.hcb), rel_err(S.hcc, oS.hcc))
    _say(f"[c4] reduced-channel job served in split precision vs the oracle {e:.2e}")
    assert e < SPLIT_TOL


def test_c5_shape_dirac(ctx, oracle_lib):
    from bsr_amd import synth_dirac as SD, Context
    w = SD.c5(nch=6)
    _say(f"[c5] ns={w.ns} ks={w.ks} ksp={w.ksp} ksq={w.ksq} mk={w.mk} nch={w.nch} records={w.n_records}")
    ctx.dirac_wave(w)
    got = ctx.dirac_get()
    ref = oracle_lib.dirac_assemble(w)
    e = max(rel_err(got[k], ref[k]) for k in ("hch", "hcp", "hp", "acf", "htarg")
            if ref[k].size and np.abs(ref[k]).max() > 0)
    _say(f"[c5] FP64 vs the oracle {e:.2e}")
    assert e < TOL
    sc = Context(0)
    try:
        sc.set_precision(2, 7)
        sc.dirac_wave(w)
        es = sc.dirac_diff(ctx)
        _say(f"[c5] int8 split (7 slices) vs FP64 {es:.2e}")
        assert 0.0 < es < SPLIT_TOL
    finally:
        sc.close()
