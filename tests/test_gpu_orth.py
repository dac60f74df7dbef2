"""GPU parity of BS_ORTH (bsrg_bs_orth, replacing
src/bsr/BSR_MAT3/bs_orth.f90:2-92) through the C ABI against the oracle:
on a matrix assembled call by call and on the H of a one-shot wave.
Tolerance 1e-12 relative (floor 1e-3 of the maximum), as for the R^k path.
"""
import numpy as np
import pytest

from tests.conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _wave(seed=31, nch=6, ncp=5):
    from bsr_amd import synth
    w = synth.make_wave(seed, 8, 40, 4.0, 12.0, nch=nch, ncp=ncp, mk=2, nb=5, ntarg=3, fill=0.3, name="orth")
    # one l for all orbitals: every bound orbital that a channel's lists do
    # not use (IBORT = 0) becomes an orthogonality condition of that channel
    w.lbs[:] = 1
    return w


def _cmp(got, ref, what):
    for name in ("hcc", "hcb"):
        a, b = getattr(got, name), getattr(ref, name)
        assert rel_err(a, b) < TOL, f"{what}.{name}: {rel_err(a, b):.3e}"


def test_bs_orth_call_by_call(ctx, oracle_lib):
    O = oracle_lib
    w = _wave()
    ctx.set_grid(w.ks, w.ns, w.ml, w.me, w.h, w.t)
    ctx.set_orbitals(w.pbs, w.iech, w.lbs, w.ibort, w.eps_ovl)
    ctx.allocate_matrix(w.nch, w.ncp, w.mk)
    ow = O.Wave(O.Grid(w.ks, w.ns, w.ml, w.me, w.h, w.t), w)
    for (jt, k), lst in sorted(w.rk_lists.items()):
        if len(lst):
            ctx.i_data(jt, k, lst)
            ow.idata(jt, k, lst)
    ctx.symmetrize_diag()
    ow.symmetrize_diag()
    before = ow.get()
    print(f"[bs_orth] call by call nch={w.nch} ncp={w.ncp}", flush=True)
    ctx.bs_orth(0)
    ow.bs_orth()
    after = ow.get()
    assert rel_err(after.hcc, before.hcc) > 1e-6   # the conditions change the matrix
    _cmp(ctx.get_matrix(), after, "call by call")


@pytest.mark.parametrize("ncp", [0, 7])
def test_bs_orth_wave(ctx, oracle_lib, ncp):
    O = oracle_lib
    w = _wave(seed=32, nch=8, ncp=ncp)
    print(f"[bs_orth] wave nch={w.nch} ncp={ncp}", flush=True)
    ctx.wave_assemble(w)
    ctx.bs_orth(1)
    _, H = O.run_wave(w, orth=True)
    _cmp(ctx.wave_get(1), H, "wave H")


def test_bs_orth_errors(oracle_lib):
    from bsr_amd import Context
    from bsr_amd.lib import BsrGpuError
    w = _wave()
    with Context(0) as c:
        c.set_grid(w.ks, w.ns, w.ml, w.me, w.h, w.t)
        c.set_orbitals(w.pbs, w.iech, w.lbs, w.ibort, w.eps_ovl)
        with pytest.raises(BsrGpuError, match="no assembled wave"):
            c.bs_orth(1)
        with pytest.raises(BsrGpuError, match="which"):
            c.bs_orth(2)
        c.allocate_matrix(w.nch + 1, w.ncp, w.mk)   # a channel with no channel orbital
        with pytest.raises(BsrGpuError, match="channel orbital"):
            c.bs_orth(0)
