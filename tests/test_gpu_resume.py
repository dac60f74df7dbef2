"""Interrupted runs (BSR_MAT4's time limit, src/bsr/BSR_MAT4/state_res.f90:
187-196; Read_matrix, read_matrix.f90:2-33): a partial wave assembled call
by call is stopped after part of its R^k lists, its matrix recorded with the
interrupt markers (bsrg_matrix_record), read back into a fresh context
(bsrg_matrix_resume) and finished there.  The resumed run gives, bit for
bit, the matrices of the uninterrupted one (the same updates in the same
order on the same values), and both match the oracle.  The file's record
layout is read back independently here.
"""
import struct

import numpy as np
import pytest

from tests.conftest import HAS_GPU, rel_err

pytestmark = pytest.mark.gpu


def _start(c, w):
    c.set_grid(w.ks, w.ns, w.ml, w.me, w.h, w.t)
    c.set_orbitals(w.pbs, w.iech, w.lbs, w.ibort, w.eps_ovl)
    c.allocate_matrix(w.nch, w.ncp, w.mk)


def _overlaps_then_scale(c, w):
    for ich in range(1, w.nch + 1):
        c.update_hl(ich, ich, 1.0)
    c.o_data(w.olist, w.eps_c)
    c.symmetrize_diag()
    c.scale(w.ec)


def _lists(w):
    return [(jt, k, w.rk_lists[(jt, k)]) for k in range(w.mk + 1) for jt in range(1, 10)
            if (jt, k) in w.rk_lists and len(w.rk_lists[(jt, k)])]


def _finish(c, w):
    for ich in range(1, w.nch + 1):
        c.update_hl(ich, ich, w.etarg[w.iptar[ich - 1] - 1] - w.ec)
    c.symmetrize_diag()
    c.acf_finish(0)
    return c.get_matrix()


def _records(path):
    out = []
    with open(path, "rb") as f:
        while True:
            h = f.read(4)
            if not h:
                return out
            n = struct.unpack("<i", h)[0]
            out.append(f.read(n))
            assert struct.unpack("<i", f.read(4))[0] == n


def test_interrupt_and_resume(oracle_lib, tmp_path):
    if not HAS_GPU:
        pytest.skip("no HIP device")
    from bsr_amd import Context, synth
    w = synth.make_wave(12, 8, 40, 4.0, 12.0, nch=5, ncp=6, mk=3, nb=3, ntarg=2, fill=0.3, name="resume")
    lists = _lists(w)
    cut = len(lists) // 2
    full, part, rest = Context(0), Context(0), Context(0)
    try:
        _start(full, w)
        _overlaps_then_scale(full, w)
        for jt, k, lst in lists:
            full.i_data(jt, k, lst)
        F = _finish(full, w)
        _start(part, w)
        _overlaps_then_scale(part, w)
        for jt, k, lst in lists[:cut]:
            part.i_data(jt, k, lst)
        path = tmp_path / "bsr_mat.int"
        part.matrix_record(path, cut, 5, 0)
        _start(rest, w)
        assert rest.matrix_resume(path) == (cut, 5, 0)
        for jt, k, lst in lists[cut:]:
            rest.i_data(jt, k, lst)
        Rz = _finish(rest, w)
    finally:
        for c in (full, part, rest):
            c.close()
    for name in ("hcc", "hcb", "hbb", "acf", "htarg", "otarg"):
        assert np.array_equal(getattr(F, name), getattr(Rz, name)), name
    _, oH = oracle_lib.run_wave(w)
    for name in ("hcc", "hcb", "acf", "htarg"):
        assert rel_err(getattr(Rz, name), getattr(oH, name)) < 1e-12, name
    # the layout: header, (i,j)+block pairs ending in (-1,-1), mk, ACF, htarg, otarg, markers
    recs = _records(path)
    assert struct.unpack("<3i", recs[0]) == (w.ns, w.nch, w.ncp)
    i = 1
    while True:
        ic, jc = struct.unpack("<2i", recs[i])
        if ic <= 0:
            break
        size = len(recs[i + 1])
        assert size == (8 * w.ns * w.ns if ic <= w.nch else (8 * w.ns if jc <= w.nch else 8))
        i += 2
    assert struct.unpack("<i", recs[i + 1])[0] == w.mk
    assert struct.unpack("<3i", recs[-1]) == (cut, 5, 0)
    assert len(recs) == i + 6
