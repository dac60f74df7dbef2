"""The list builder feeding the GPU assembly: a seeded synthetic int_bnk
bank (bsr_amd/synth_bank.py), written to a file and read back by the
builder, expanded over the states of its partial wave (icase 5 R^k lists,
icase 11 overlap list), and the lists assembled by bsrg_wave_assemble into
S and H; both against the oracle's assembly of the same lists (1e-12
relative per element, floor 1e-3 of the array maximum).  The builder itself
is pinned bit for bit against the reference's routines by
tests/test_host_state_res.py.
"""
import numpy as np
import pytest

from tests.conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.mark.parametrize("seed,nch,ncp,blocks", [(1, 8, 3, "default"), (5, 12, 4, "small")])
def test_bank_to_matrices(ctx, oracle_lib, tmp_path, seed, nch, ncp, blocks):
    from bsr_amd import synth_bank as SB
    O = oracle_lib
    b = SB.make_bank(seed, nch=nch, ncp=ncp, mk=3, nb=6)
    path = tmp_path / "int_bnk.tst"
    b.write(path)
    npol = max(int(b.wave.lbs.max()), b.wave.mk)
    params = {} if blocks == "default" else dict(nb=(npol + 2) * 9 + 3, mb=3, kb=2)
    lb = SB.builder_for(b, path=str(path), **params)
    lb.run(5, b.wave.mk)
    rk = lb.lists()
    lb.run(11, b.wave.mk)
    ov = lb.lists()
    w = SB.wave_from_lists(b, rk, ov)
    print(f"[bank] seed={seed} records={len(b.c)} R^k list records={len(rk[0])} in {rk[0].max() + 1} lists, "
          f"overlap records={len(ov[0])}", flush=True)
    assert len(rk[0]) > 0 and len(ov[0]) > 0
    ctx.wave_assemble(w)
    S, H = ctx.wave_get(0), ctx.wave_get(1)
    oS, oH = O.run_wave(w)
    e = 0.0
    for got, ref in ((S, oS), (H, oH)):
        for name in ("hcc", "hcb", "hbb", "acf", "htarg", "otarg"):
            r = np.asarray(getattr(ref, name))
            g = np.asarray(getattr(got, name))
            if r.size and np.abs(r).max() > 0:
                e = max(e, rel_err(g, r))
            else:
                assert not g.any(), name
    print(f"[bank] S and H vs the oracle {e:.2e}", flush=True)
    assert e < TOL
