"""R^k lists decoded on the device (bsrg_lists_upload, then
bsrg_wave_assemble_resident) against the host decoder (bsrg_wave_assemble)
on the same wave: S, H and the host scalars (HB, target matrices, ACF) are
equal bit for bit -- both decoders add every output's terms in list order
-- on waves with every record structure, with and without perturbers, at
the C2 shape, repeated and after other waves; a record the host decoder
rejects gives the same status and message on the device."""
import copy

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("hcc", "hcb", "hbb", "acf", "htarg", "otarg")


def _wave(name):
    from bsr_amd import synth as S
    if name == "all_structures":
        return S.make_wave(11, 8, 40, 4.0, 12.0, nch=6, ncp=7, mk=3, nb=4, ntarg=3, fill=0.3, name="small")
    if name == "c1":
        return S.c1()
    if name == "perturbers":
        return S.make_wave(31, 8, 120, 18.0, 40.0, nch=10, ncp=30, mk=6, nb=5, ntarg=4, fill=0.1, name="pert")
    if name == "c2":
        return S.c2()
    raise KeyError(name)


def _get(ctx):
    return ctx.wave_get(0), ctx.wave_get(1)


def _equal(host, dev, what):
    for m, (a, b) in enumerate(zip(host, dev)):
        for f in FIELDS:
            x, y = getattr(a, f), getattr(b, f)
            assert x.shape == y.shape, f"{what} {'SH'[m]}.{f}: shapes {x.shape} {y.shape}"
            if not np.array_equal(x, y):
                d = np.abs(x - y).max() if x.size else 0.0
                raise AssertionError(f"{what} {'SH'[m]}.{f}: not bit-equal, max |diff| {d:.3e}")


@pytest.mark.parametrize("name", ["all_structures", "c1", "perturbers", "c2"])
def test_resident_equals_host(ctx, name):
    w = _wave(name)
    print(f"[resident] {name}: nhm={w.nhm} records={w.n_records()}", flush=True)
    ctx.wave_assemble(w)
    host = _get(ctx)
    hc = ctx.wave_term_counts(1)
    ctx.lists_upload(w)
    ctx.wave_assemble_resident(w)
    _equal(host, _get(ctx), name)
    dc = ctx.wave_term_counts(1)
    for key in ("blocks", "band_blocks", "band_terms", "rank1_blocks", "rank1_terms", "vec_positions", "vec_terms"):
        assert dc[key] == hc[key], (key, dc[key], hc[key])
    # again: the same result
    ctx.wave_assemble_resident(w)
    _equal(host, _get(ctx), name + " (repeat)")


@pytest.mark.parametrize("precision", [(2, 7), (1, 6)], ids=["int8x7", "bf16x6"])
def test_resident_split_precision(ctx, precision):
    """The exchange product in split precision (int8 or bf16 slices): the
    same coefficients from either decoder give the same bits."""
    w = _wave("perturbers")
    try:
        ctx.set_precision(*precision)
        ctx.wave_assemble(w)
        host = _get(ctx)
        ctx.lists_upload(w)
        ctx.wave_assemble_resident(w)
        _equal(host, _get(ctx), f"split {precision}")
    finally:
        ctx.set_precision(0, 0)


def test_resident_after_other_waves(ctx):
    """Lists stay resident across waves of other shapes assembled on the host
    in between; the GET_V vectors of each path are their own."""
    a, b = _wave("all_structures"), _wave("perturbers")
    ctx.wave_assemble(a)
    ha = _get(ctx)
    ctx.wave_assemble(b)
    hb = _get(ctx)
    ctx.lists_upload(a)
    ctx.wave_assemble_resident(a)
    _equal(ha, _get(ctx), "a")
    ctx.wave_assemble(b)
    _equal(hb, _get(ctx), "b on the host after a resident wave")
    ctx.wave_assemble_resident(a)
    _equal(ha, _get(ctx), "a again")
    ctx.lists_upload(b)
    ctx.wave_assemble_resident(b)
    _equal(hb, _get(ctx), "b resident")


def _mutated(w, jtype, field, value):
    """A copy of w whose first record of structure jtype has field (k1..k4)
    set to value."""
    jt, kp, k1, k2, k3, k4, c = (np.array(x) for x in w.concat())
    r = int(np.flatnonzero(jt == jtype)[0])
    arrays = {"k1": k1, "k2": k2, "k3": k3, "k4": k4}
    arrays[field][r] = value
    w2 = copy.copy(w)
    w2.__dict__ = dict(w.__dict__)
    w2.__dict__["_packed"] = (jt, kp, k1, k2, k3, k4, c)
    return w2


def test_resident_rejects_as_host(ctx):
    from bsr_amd import synth as S
    from bsr_amd.lib import BsrGpuError
    w = _wave("all_structures")
    cases = [
        (6, "k1", (w.nch + 1) * S.IBI + 1, -10),   # channel index outside the wave
        (1, "k3", -(w.ncp + 1), -12),              # perturber index outside the wave (UPDATE_HB)
        (2, "k2", w.nb + 1, -31),                  # BAV with a channel orbital
    ]
    for jtype, field, value, code in cases:
        bad = _mutated(w, jtype, field, value)
        if jtype == 1:   # a UPDATE_HB record: both perturber indices positive
            jt, kp, k1, k2, k3, k4, c = bad.concat()
            r = int(np.flatnonzero(jt == 1)[0])
            k4[r] = -1
        with pytest.raises(BsrGpuError) as eh:
            ctx.wave_assemble(bad)
        ctx.lists_upload(bad)
        with pytest.raises(BsrGpuError) as ed:
            ctx.wave_assemble_resident(bad)
        print(f"[resident] jtype {jtype}: host {eh.value} device {ed.value}", flush=True)
        assert eh.value.code == ed.value.code == code
        assert str(eh.value) == str(ed.value)
    # the context stays usable
    ctx.wave_assemble(w)
    host = _get(ctx)
    ctx.lists_upload(w)
    ctx.wave_assemble_resident(w)
    _equal(host, _get(ctx), "after rejections")
