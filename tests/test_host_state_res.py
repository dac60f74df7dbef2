"""The list builder (bsrg_lb_*, bsr_amd/csrc/state_res.cpp) against the
reference's own routines (no GPU).

The reference build oracle/_ref/liblb_ref.so (ref.mk; the reference's
Read_dets with its int_bnk header readers, Decode_INT, Ifind_channel,
Jsym_int, Det_fact, Check_det, VDET, DET and Idef_itype, compiled from its
sources) is walked over the records in State_res's order by
oracle/ref_lb_driver.f90; the lists go through the reference's Add_cdata and
Merge_CDATA (oracle/state_res.py).  On seeded synthetic banks
(bsr_amd/synth_bank.py) the builder must give, bit for bit:

  * the bank as the reference reads it (records, determinants, factors)
  * every coefficient State_res passes to Add_coef, in order, with its
    structure, multipole and key (icase 5 R^k, 6 L, 11 overlaps)
  * the lists Gen_matrix hands on, with the default blocks and with blocks
    so small that full sets force generation in the middle of the walk

and its error paths must give the reference's stops as statuses.
"""

import numpy as np
import pytest

from bsr_amd import lib as L
from bsr_amd import synth_bank as SB
from oracle import refbind as R

needs_ref = pytest.mark.skipif(not (R.lb_available() and R.available()),
                               reason="reference builds oracle/_ref/liblb_ref.so and libbsr_ref.so absent")

SEEDS = [(0, 8, 3), (1, 8, 3), (5, 10, 4), (7, 12, 0)]


def _bank(seed, nch, ncp):
    return SB.make_bank(seed, nch=nch, ncp=ncp, mk=3, nb=6)


def test_incode_matches_decode_layout():
    """Incode_INT's packing (the writer's) against its inverse written out."""
    rng = np.random.default_rng(1)
    for _ in range(200):
        met, k = int(rng.integers(0, 16)), int(rng.integers(0, 60))
        i = [int(x) for x in rng.integers(0, 32, 4)]
        v = SB.incode_int(met, k, *i)
        assert v % 32 == i[0] and v // 32 % 32 == i[1] and v // 32 ** 2 % 32 == i[2] and v // 32 ** 3 % 32 == i[3]
        assert v // 32 ** 4 % 16 == met and v // (32 ** 4 * 16) == k


@needs_ref
def test_decode_and_jsym_reference():
    """The reference's Decode_INT inverts the writer's packing; Jsym_int's
    canonical forms on every case (the builder's use is checked through the
    expansion below)."""
    rng = np.random.default_rng(2)
    n = 300
    met = rng.integers(1, 12, n)
    k = rng.integers(0, 20, n)
    q = rng.integers(1, 32, (4, n))
    ints = np.array([SB.incode_int(int(met[j]), int(k[j]), *(int(x) for x in q[:, j])) for j in range(n)])
    out = R.lb_decode(ints)
    assert np.array_equal(out[0], met) and np.array_equal(out[1], k) and np.array_equal(out[2:], q)
    js = R.lb_jsym(met, q)
    for j in range(n):
        a = q[:, j]
        b = js[:, j]
        # the canonical form is one of the integral's symmetric images with its smallest index first where allowed
        assert sorted(a) == sorted(b)
        if met[j] in (4, 5):
            assert b[0] == a.min()


@needs_ref
@pytest.mark.parametrize("seed,nch,ncp", SEEDS)
def test_bank_reader_vs_reference(tmp_path, seed, nch, ncp):
    b = _bank(seed, nch, ncp)
    path = tmp_path / "int_bnk.tst"
    b.write(path)
    (c, ijt, intb, idf), dets = R.lb_read(str(tmp_path))
    lb = SB.builder_for(b, path=str(path))
    got = lb.bank()
    for key, ref in zip(("c", "ijt", "intb", "idf"), (c, ijt, intb, idf)):
        assert np.array_equal(got[key], ref), key
    for key, ref in zip(("kpd", "ipd", "npd", "kpf", "ipf", "npf"), dets):
        assert np.array_equal(got[key], ref), key
    assert np.array_equal(got["c"], b.c) and np.array_equal(got["intb"], b.intb)


@needs_ref
@pytest.mark.parametrize("seed,nch,ncp", SEEDS)
@pytest.mark.parametrize("icase", [5, 6, 11])
def test_expansion_vs_reference(seed, nch, ncp, icase):
    b = _bank(seed, nch, ncp)
    lb = SB.builder_for(b)
    lb.run(icase, b.wave.mk)
    got = lb.expansion()
    ref = R.lb_expand(b, icase, b.wave.mk)
    assert len(ref[0]) > 0
    for q, (x, y) in enumerate(zip(got, ref)):
        assert np.array_equal(x, y), q


@needs_ref
def test_expansion_covers_the_structures():
    """Over the seeds, R^k records reach structures 1, 3-5, 7 and 9 (the
    canonical forms of Jsym_int put a bound orbital first, so 2, 6 and 8 need
    a channel orbital of lower index), L records 4 and 6, overlaps 1."""
    seen = {5: set(), 6: set(), 11: set()}
    for seed, nch, ncp in SEEDS:
        b = _bank(seed, nch, ncp)
        for icase in seen:
            lb = SB.builder_for(b)
            lb.run(icase, b.wave.mk)
            seen[icase] |= set(lb.expansion()[0].tolist())
    assert {1, 3, 4, 5, 7, 9} <= seen[5]
    assert {4, 6} <= seen[6]
    assert seen[11] == {1}


@needs_ref
@pytest.mark.parametrize("seed,nch,ncp", SEEDS[:3])
@pytest.mark.parametrize("blocks", ["default", "small"])
def test_lists_vs_reference(seed, nch, ncp, blocks):
    from oracle import state_res as OS
    b = _bank(seed, nch, ncp)
    npol = max(int(b.wave.lbs.max()), b.wave.mk)
    params = {} if blocks == "default" else dict(nb=(npol + 2) * 9 + 3, mb=3, kb=2)
    for icase in (5, 11):
        lb = SB.builder_for(b, **params)
        sz = lb.run(icase, b.wave.mk)
        got = lb.lists()
        ref = OS.build_lists(R.lb_expand(b, icase, b.wave.mk), npol, **params)
        for q, (x, y) in enumerate(zip(got, ref)):
            assert np.array_equal(x, y), (icase, q)
        if blocks == "small" and icase == 5:
            assert sz[2] > len(set(zip(got[1].tolist(), got[2].tolist())))   # generation in mid-walk


def test_errors():
    b = _bank(0, 8, 3)
    lb = SB.builder_for(b)
    with pytest.raises(L.BsrGpuError) as e:
        lb.run(8, b.wave.mk)
    assert e.value.code == -44
    # a channel-channel integral over bound orbitals with no overlap factor: no structure fits
    bad = SB.make_bank(0, nch=8, ncp=3, mk=3, nb=6)
    cc = int(np.nonzero(bad.it1 > 0)[0][0]) + 1
    bad.c = np.array([0.5])
    bad.ijt = np.array([cc * SB.IBC + cc], np.int32)
    bad.intb = np.array([SB.incode_int(5, 0, 1, 2, 1, 2)], np.int32)
    bad.idf = np.array([0], np.int32)
    lb = SB.builder_for(bad)
    with pytest.raises(L.BsrGpuError) as e:
        lb.run(5, 3)
    assert e.value.code == -48 and "Idef_itype=0" in str(e.value)


def test_bank_file_errors(tmp_path):
    from bsr_amd.listbuild import ListBuilder
    lb = ListBuilder()
    with pytest.raises(L.BsrGpuError) as e:
        lb.read_bank(str(tmp_path / "missing"))
    assert e.value.code == -41
    p = tmp_path / "short"
    p.write_bytes(b"\x08\x00\x00\x00\x01\x00")
    with pytest.raises(L.BsrGpuError) as e:
        lb.read_bank(str(p))
    assert e.value.code == -40
