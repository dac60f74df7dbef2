"""Dirac partial waves on the GPU (R_data over the four integral families,
through the C ABI) against the oracle (oracle/dbsr_oracle.c, which equals
the reference's own R_data bit for bit on the fixture waves, see
test_oracle_dirac.py).  Every element of hch, hcp, hp, acf and htarg within
1e-12 relative (floor 1e-3 of the array's max); diagonal channel blocks
exactly symmetric.  Cases: perturbers or none, unequal (ksp < ksq) and equal
component orders, several multipoles; a second wave on the same context
(smaller, buffers reused); the rejected inputs."""
import numpy as np
import pytest

from bsr_amd import synth_dirac as SD
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12
KEYS = ("hch", "hcp", "hp", "acf", "htarg")

CASES = [
    ("small_pert", lambda: SD.small(11, 3)),
    ("small_nopert", lambda: SD.small(12, 0)),
    ("ks9_pert", lambda: SD.make_dirac_wave(21, 9, 40, 8, 9, 54.0, 30.0, 4, 6, 2, 4, 0.15, name="ks9")),
    ("ks9_equal_orders", lambda: SD.make_dirac_wave(22, 9, 36, 9, 9, 30.0, 25.0, 5, 4, 0, 3, 0.2)),
]


def _check(ctx, O, w, name):
    print(f"[dirac] case={name} ns={w.ns} ks={w.ks} ksp={w.ksp} ksq={w.ksq} nch={w.nch} npert={w.npert} "
          f"nb={w.nb} mk={w.mk} records={w.n_records}", flush=True)
    ctx.dirac_wave(w)
    got = ctx.dirac_get()
    ref = O.dirac_assemble(w)
    for key in KEYS:
        assert got[key].shape == ref[key].shape, key
        e = rel_err(got[key], ref[key])
        assert e < TOL, (key, e)
    for i in range(w.nch):
        b = got["hch"][:, :, i]
        assert np.array_equal(b, b.T), i
    return got


@pytest.mark.parametrize("case", CASES, ids=lambda c: c[0])
def test_dirac_wave(ctx, oracle_lib, case):
    name, make = case
    _check(ctx, oracle_lib, make(), name)


@pytest.mark.parametrize("engine", [0, 2, 4])
def test_dirac_fp64_engines(oracle_lib, engine):
    """Every FP64 exchange engine against the oracle: the library GEMM (0)
    and the MFMA GEMM kernel (2) into the quadrant-major buffer, then the
    term pass into the output layout; the sparse product (4, the default)
    straight into the output layout, then the term pass in place (quadrants
    without terms left as the product wrote them)."""
    from bsr_amd import Context
    from tests.conftest import HAS_GPU
    if not HAS_GPU:
        pytest.skip("no HIP device")
    c = Context(0)
    try:
        c.set_gemm_engine(engine)
        for name, make in CASES[:3]:
            _check(c, oracle_lib, make(), f"{name}_engine{engine}")
            assert c.dirac_engine() == engine
    finally:
        c.close()


def test_dirac_reuse(ctx, oracle_lib):
    """A larger wave, then a smaller one with other orbitals, then the first
    again: resident bases are rebuilt or kept as the orbitals require."""
    big = CASES[2][1]()
    small = CASES[0][1]()
    _check(ctx, oracle_lib, big, "reuse_big")
    _check(ctx, oracle_lib, small, "reuse_small")
    _check(ctx, oracle_lib, big, "reuse_big_again")


def test_dirac_rejects(ctx):
    from bsr_amd.lib import BsrGpuError as BsrgError
    w = SD.small(11, 3)
    ctx.dbs_set_grid(w.ks, w.ns, w.ksp, w.ksq, w.t)
    ctx.dirac_set_orbitals(w.pq, w.ipbs, w.i_ch, w.j_ch, w.v_ch)
    ibo = SD.IBO
    # exchange record with the smaller channel first
    with pytest.raises(BsrgError) as e:
        ctx.dirac_assemble(w.nch, w.npert, w.mk, w.ntarg, w.iptar, [4], [0], [1 * ibo + 2], [1], [2], [0], [0.5])
    assert e.value.code == -25
    # overlap without a <.|p> vector (KBORT = 0): bound orbital 1 with itself
    with pytest.raises(BsrgError) as e:
        ctx.dirac_assemble(w.nch, w.npert, w.mk, w.ntarg, w.iptar, [2], [0], [1 * ibo + 1], [1], [1],
                           [1 * ibo + 1], [0.5])
    assert e.value.code == -19
    # density of a channel orbital
    with pytest.raises(BsrgError) as e:
        ctx.dirac_assemble(w.nch, w.npert, w.mk, w.ntarg, w.iptar, [3], [0], [2 * ibo + 1], [1], [w.nb + 1],
                           [0], [0.5])
    assert e.value.code == -31


@pytest.mark.parametrize("mode,nslices", [(2, 7), (2, 8), (1, 7), (1, 8)])
def test_dirac_split_precision(ctx, mode, nslices):
    """The exchange GEMMs of the four families in split precision (int8 or
    bf16 Ozaki slices, FP64 recombination): every element within 1e-10 of
    the FP64 result of the same wave (floor 1e-3 of the largest matrix
    element).  Dirac waves need 7 slices for that (6 slices measure 3-4e-10:
    the rows of the exchange matrices span all multipoles and bound pairs)."""
    from bsr_amd import Context
    w = CASES[2][1]()
    print(f"[dirac-split] mode={mode} S={nslices} ns={w.ns} nch={w.nch}", flush=True)
    ctx.dirac_wave(w)
    sc = Context(0)
    try:
        sc.set_precision(mode, nslices)
        sc.dirac_wave(w)
        assert sc.dirac_engine() == 10 * mode + nslices
        err = sc.dirac_diff(ctx)
        print(f"[dirac-split] max rel element error vs FP64 {err:.2e}", flush=True)
        assert err < 1e-10
        assert err > 0.0  # the split path really ran
    finally:
        sc.close()
