"""Dirac B-spline bands on the GPU (through the C ABI) against the oracle:
moments of both bases, the four R^k band families, and the convolutions,
on the reference's DBS test grid (ksp=8, ksq=9) and grids with equal
orders.  Tolerance 1e-12 relative; bands are non-negative."""
import numpy as np
import pytest

from tests.conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _grids():
    import os
    gv = np.load(os.path.join(os.path.dirname(__file__), "golden", "dbs_vectors.npz"))
    out = []
    for name in ("d1", "d2"):
        ks, ns, ksp, ksq = (int(v) for v in gv[f"{name}_ints"])
        out.append((name, ks, ns, ksp, ksq, gv[f"{name}_t"]))
    t = np.concatenate([np.zeros(9), np.geomspace(0.005, 30.0, 60), np.full(8, 30.0)])
    out.append(("eq9", 9, len(t) - 9, 9, 9, t))
    return out


@pytest.mark.parametrize("case", _grids(), ids=lambda c: c[0])
def test_dbs_bands(ctx, oracle_lib, case):
    O = oracle_lib
    name, ks, ns, ksp, ksq, t = case
    print(f"[dbs] grid={name} ks={ks} ns={ns} ksp={ksp} ksq={ksq}", flush=True)
    ctx.dbs_set_grid(ks, ns, ksp, ksq, t)
    og = O.DbsGrid(ks, ns, ksp, ksq, t)
    for q in (0, 1):
        kb = ksq if q else ksp
        for k in (0, 3):
            a, b = ctx.dbs_moments(q, k)
            assert rel_err(a, og.moments(q, k)[: kb * (kb + 1) // 2], 1e-14) < TOL
            assert rel_err(b, og.moments(q, -(k + 1))[: kb * (kb + 1) // 2], 1e-14) < TOL
    for f in range(4):
        for k in (0, 2, 5):
            got, ref = ctx.dbs_mrk(f, k), og.mrk(f, k)
            assert np.all(got >= 0.0)
            assert np.array_equal(got != 0, ref != 0), (f, k)
            assert rel_err(got, ref, 1e-14) < TOL, (f, k)


def test_dbs_convolutions(ctx, oracle_lib):
    O = oracle_lib
    name, ks, ns, ksp, ksq, t = _grids()[0]
    ctx.dbs_set_grid(ks, ns, ksp, ksq, t)
    og = O.DbsGrid(ks, ns, ksp, ksq, t)
    rng = np.random.default_rng(5)
    for f in range(4):
        rkb = og.mrk(f, 3)
        for icase in (1, 2, 3, 4):
            d = rng.standard_normal((ns, ks) if icase <= 2 else (ns, ns))
            d = np.asfortranarray(d)
            got = ctx.dbs_convol(f, 3, d, icase)
            ref = O.convol(ns, ks, d, icase, rkb)
            assert rel_err(got, ref) < TOL, (f, icase)
