"""GPU parity beyond the synthetic defaults: more bound orbitals than one
exchange-basis tile (the reference's I_data/R_data take any nbf,
src/bsr/BSR_MAT3/I_data.f90:192-216), BS_ORTH with many conditions per
channel and with work arrays too large for LDS (bs_orth.f90:36-92 has no
limit), on the knot grid of the reference's e+C example
(examples/bsr/C_LS/initial files/knot.dat: ks=8, ns=121, z=6, h=0.25,
hmax=0.5, rmax=50; its target file lists nwt = 18 orbitals).

Tolerance: 1e-12 relative to the oracle per element (floor 1e-3 of the
array maximum), as for every FP64 path.
"""
import numpy as np
import pytest

from tests.conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _ec_grid():
    from bsr_amd import grid as G
    g = G.recipe_grid(0, 8, 6.0, 0.25, 0.5, 50.0)
    assert g.ns == 121
    return g


def _cmp(got, ref, what, names=("hcc", "hcb", "hbb", "acf", "htarg", "otarg")):
    for name in names:
        a, b = getattr(got, name), getattr(ref, name)
        if b.size and np.abs(b).max() > 0:
            e = rel_err(a, b)
            assert e < TOL, f"{what}.{name}: {e:.3e}"
        else:
            assert np.abs(a).max() == 0.0, f"{what}.{name} should be zero"


@pytest.mark.parametrize("nb", [9, 18, 32])
def test_wave_many_bound_orbitals(ctx, oracle_lib, nb):
    """A fused wave with nb bound orbitals on the e+C grid (exchange bases
    in tiles of 8 orbitals: 9 leaves a one-orbital tile, 18 a two-orbital
    one, 32 four full tiles) against the oracle."""
    O = oracle_lib
    from bsr_amd import synth
    w = synth.make_wave(70 + nb, 8, 0, 0, 0, nch=5, ncp=4, mk=3, nb=nb, ntarg=3, fill=0.04, name="eC", grid=_ec_grid())
    print(f"[general] e+C grid ns={w.ns} nb={nb} nch={w.nch} records={w.n_records()}", flush=True)
    ctx.wave_assemble(w)
    S, H = ctx.wave_get(0), ctx.wave_get(1)
    oS, oH = O.run_wave(w)
    _cmp(S, oS, "S", ("hcc", "hcb", "hbb", "otarg"))
    _cmp(H, oH, "H")


def test_call_by_call_18_orbitals(ctx, oracle_lib):
    """I_data call by call (exchange through the same tiled bases) with 18
    bound orbitals, as the reference's e+C deck has."""
    O = oracle_lib
    from bsr_amd import synth
    g = _ec_grid()
    w = synth.make_wave(88, 8, 0, 0, 0, nch=4, ncp=3, mk=2, nb=18, ntarg=3, fill=0.05, name="eC", grid=g)
    ctx.set_grid(w.ks, w.ns, w.ml, w.me, w.h, w.t)
    ctx.set_orbitals(w.pbs, w.iech, w.lbs, w.ibort, w.eps_ovl)
    ctx.allocate_matrix(w.nch, w.ncp, w.mk)
    ow = O.Wave(O.Grid(w.ks, w.ns, w.ml, w.me, w.h, w.t), w)
    for k in range(w.mk + 1):
        for jt in range(1, 10):
            lst = w.rk_lists.get((jt, k))
            if lst is not None and len(lst):
                ctx.i_data(jt, k, lst)
                ow.idata(jt, k, lst)
    print("[general] call by call nb=18", flush=True)
    _cmp(ctx.get_matrix(), ow.get(), "call by call")


def test_gen_lval_closed_core_beyond_one_tile(ctx, oracle_lib):
    """Gen_Lval with a closed core of 3 orbitals and 12 bound orbitals."""
    O = oracle_lib
    from bsr_amd import synth
    w = synth.make_wave(90, 8, 60, 6.0, 20.0, nch=4, ncp=3, mk=2, nb=12, ntarg=3, fill=0.1, name="lval")
    ctx.set_grid(w.ks, w.ns, w.ml, w.me, w.h, w.t)
    ctx.set_orbitals(w.pbs, w.iech, w.lbs, w.ibort, w.eps_ovl)
    ctx.gen_lval(6.0, 3)
    ow = O.Wave(O.Grid(w.ks, w.ns, w.ml, w.me, w.h, w.t), w)
    ow.gen_lval(6.0, 3)
    mlh, hl, li, lv = ctx.lval()
    mr, hr, lr, vr = ow.lval()
    print(f"[general] gen_lval nb=12 kclosd=3 mlh={mlh}", flush=True)
    assert mlh == mr
    assert rel_err(hl, hr) < TOL
    assert rel_err(li, lr) < TOL
    assert rel_err(lv, vr) < TOL


@pytest.mark.parametrize("ns,nb", [(121, 24), (240, 20)], ids=["lds_chunks", "global_scratch"])
def test_bs_orth_many_conditions(ctx, oracle_lib, ns, nb):
    """Every orbital of one l: each channel gets the bound orbitals its lists
    do not exempt as conditions (more than 16 here).  At ns=121 the work
    arrays fit LDS and the conditions run in chunks; at ns=240 they go to
    global scratch."""
    O = oracle_lib
    from bsr_amd import synth
    if ns == 121:
        w = synth.make_wave(95, 8, 0, 0, 0, nch=4, ncp=3, mk=1, nb=nb, ntarg=3, fill=0.05, name="orth", grid=_ec_grid())
    else:
        w = synth.make_wave(96, 8, ns, 6.0, 30.0, nch=3, ncp=2, mk=1, nb=nb, ntarg=3, fill=0.05, name="orth")
    w.lbs[:] = 0
    nz = [int(np.sum(w.ibort[nb + c, :nb] == 0)) for c in range(w.nch)]
    print(f"[general] bs_orth ns={w.ns} nb={nb} conditions per channel {nz}", flush=True)
    assert max(nz) > 16
    ctx.wave_assemble(w)
    ctx.bs_orth(1)
    _, H = O.run_wave(w, orth=True)
    _cmp(ctx.wave_get(1), H, "orth H", ("hcc", "hcb"))


def test_dirac_wave_many_bound_orbitals(ctx, oracle_lib):
    """A Dirac wave (four families) with 11 bound orbitals."""
    from bsr_amd import synth_dirac as SD
    w = SD.make_dirac_wave(31, 6, 30, 5, 6, 20.0, 20.0, 11, 4, 2, 2, 0.08, name="dirac_nb11")
    print(f"[general] dirac nb={w.nb} records={w.n_records}", flush=True)
    ctx.dirac_wave(w)
    got = ctx.dirac_get()
    ref = oracle_lib.dirac_assemble(w)
    for key in ("hch", "hcp", "hp", "acf", "htarg"):
        assert rel_err(got[key], ref[key]) < TOL, key
