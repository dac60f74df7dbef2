"""GPU parity of the one-electron operator with the core (bsrg_gen_lval,
replacing Gen_Lval / INT_core of src/bsr/BSR_MAT3/l_core.f90) and of the
L-integral lists (bsrg_ldata, replacing L_data.f90) through the C ABI
against the oracle.

Tolerance: 1e-12 relative to the oracle (floor 1e-3 of the maximum for
elements that cancel), as for the R^k path.
"""
import numpy as np
import pytest

from tests.conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _small_wave():
    from bsr_amd import synth
    return synth.make_wave(11, 8, 40, 4.0, 12.0, nch=6, ncp=7, mk=3, nb=4, ntarg=3, fill=0.3, name="small")


def _setup(ctx, w):
    ctx.set_grid(w.ks, w.ns, w.ml, w.me, w.h, w.t)
    ctx.set_orbitals(w.pbs, w.iech, w.lbs, w.ibort, w.eps_ovl)


def _cmp_mats(got, ref, what):
    for name in ("hcc", "hcb", "hbb", "htarg", "otarg"):
        a, b = getattr(got, name), getattr(ref, name)
        if b.size and np.abs(b).max() > 0:
            e = rel_err(a, b)
            assert e < TOL, f"{what}.{name}: {e:.3e}"
        else:
            assert np.abs(a).max() == 0.0, f"{what}.{name} should be zero"


@pytest.mark.parametrize("kclosd", [0, 2])
def test_gen_lval(ctx, oracle_lib, kclosd):
    O = oracle_lib
    w = _small_wave()
    _setup(ctx, w)
    print(f"[gen_lval] kclosd={kclosd} ns={w.ns}", flush=True)
    ctx.gen_lval(4.0, kclosd)
    ow = O.Wave(O.Grid(w.ks, w.ns, w.ml, w.me, w.h, w.t), w)
    ow.gen_lval(4.0, kclosd)
    mlh, hl, li, lv = ctx.lval()
    mr, hr, lr, vr = ow.lval()
    assert mlh == mr
    for l in range(mlh + 1):
        if np.abs(hr[:, :, l]).max() > 0:
            assert rel_err(hl[:, :, l], hr[:, :, l]) < TOL, l
        else:
            assert not hl[:, :, l].any(), l
    assert rel_err(li, lr) < TOL
    assert rel_err(lv, vr) < TOL


@pytest.mark.parametrize("iitar", [0, 2])
@pytest.mark.parametrize("jtype", range(1, 10))
def test_ldata_each_structure(ctx, oracle_lib, jtype, iitar):
    O = oracle_lib
    from bsr_amd import synth
    w = _small_wave()
    _setup(ctx, w)
    ctx.allocate_matrix(w.nch, w.ncp, w.mk)
    ctx.gen_lval(4.0, 2)
    ow = O.Wave(O.Grid(w.ks, w.ns, w.ml, w.me, w.h, w.t), w)
    ow.gen_lval(4.0, 2)
    n = 0
    for (jt, l), lst in sorted(synth.l_lists(w, 5, fill=0.3).items()):
        if jt != jtype or len(lst) == 0:
            continue
        print(f"[ldata] jtype={jt} l={l} n={len(lst)} iitar={iitar}", flush=True)
        ctx.l_data(jt, l, lst, iitar=iitar)
        ow.ldata(jt, l, lst, iitar=iitar)
        n += len(lst)
    assert n > 0
    _cmp_mats(ctx.get_matrix(), ow.get(), f"jtype{jtype}")


def test_l_then_rk_on_one_matrix(ctx, oracle_lib):
    """SUB1's order: L-integral lists, then the R^k lists, into one H."""
    O = oracle_lib
    from bsr_amd import synth
    w = _small_wave()
    _setup(ctx, w)
    ctx.allocate_matrix(w.nch, w.ncp, w.mk)
    ctx.gen_lval(4.0, 1)
    ow = O.Wave(O.Grid(w.ks, w.ns, w.ml, w.me, w.h, w.t), w)
    ow.gen_lval(4.0, 1)
    for (jt, l), lst in sorted(synth.l_lists(w, 9, fill=0.2).items(), key=lambda kv: (kv[0][1], kv[0][0])):
        if len(lst):
            ctx.l_data(jt, l, lst, iitar=1)
            ow.ldata(jt, l, lst, iitar=1)
    for (jt, k), lst in sorted(w.rk_lists.items()):
        if len(lst):
            ctx.i_data(jt, k, lst)
            ow.idata(jt, k, lst)
    _cmp_mats(ctx.get_matrix(), ow.get(), "L+Rk")


def test_ldata_errors(oracle_lib):
    from bsr_amd import Context, synth
    from bsr_amd.lib import BsrGpuError
    w = _small_wave()
    lst = synth.RecordList.build([1 * synth.IBI + 1], [0], [0], [0], [1.0])
    with Context(0) as c:
        _setup(c, w)
        c.allocate_matrix(w.nch, w.ncp, w.mk)
        with pytest.raises(BsrGpuError, match="Gen_Lval"):
            c.l_data(6, 0, lst)
        c.gen_lval(4.0, 0)
        with pytest.raises(BsrGpuError, match="l outside"):
            c.l_data(6, 99, lst)
        with pytest.raises(BsrGpuError, match="itype"):
            c.l_data(10, 0, lst)
        with pytest.raises(BsrGpuError, match="unknown structure"):
            c.l_data(1, 0, synth.RecordList.build([1 * synth.IBI + 1], [0], [-3], [0], [1.0]))
        with pytest.raises(BsrGpuError, match="kclosd"):
            c.gen_lval(4.0, 99)
        # new orbitals invalidate the operator
        c.set_orbitals(w.pbs * 1.001, w.iech, w.lbs, w.ibort, w.eps_ovl)
        with pytest.raises(BsrGpuError, match="Gen_Lval"):
            c.l_data(6, 0, lst)
