"""Host-side checks that need no GPU: the C ABI loads and exports what the
header declares, the host grid set-up matches the oracle, the workload
generator is deterministic, and the bench refuses more than one GPU."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    from bsr_amd import lib as L
    header = open(os.path.join(ROOT, "include", "bsr_gpu.h")).read()
    names = sorted(set(re.findall(r"\b(bsrg_\w+)\s*\(", header)))
    assert len(names) >= 20
    lib = L.lib()
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(L.SIGNATURES), "python binding and header disagree"
    assert lib.bsrg_abi_version() == 1


@pytest.mark.parametrize("args", [(0, 8, 2.0, 0.125, 160.0, 160.0), (1, 8, 6.0, 0.2, 0.9, 20.0),
                                  (1, 9, 54.0, 0.1, 0.5, 30.0), (1, 8, 18.0, 0.05, 0.45, 40.0)])
def test_knots_match_oracle(args):
    from bsr_amd import grid as G
    kind = args[0]
    a = G.mkgrid(*args)
    b = (O.mkgrid0 if kind == 0 else O.mkgrid1)(*args[1:])
    assert a[1:] == b[1:]
    assert np.array_equal(a[0], b[0])


def test_tables_match_oracle():
    from bsr_amd import grid as G
    g = G.fixed_ns_grid(8, 60, 6.0, 20.0, 0.05)
    gr, grw, bsp, sb = G.tables(g)
    og = O.Grid(g.ks, g.ns, g.ml, g.me, g.h, g.t)
    ogr, ogrw, obsp, osb = og.tables()
    assert np.max(np.abs(gr - ogr)) <= 1e-15 * np.abs(ogr).max()
    assert np.max(np.abs(grw - ogrw)) <= 1e-15 * np.abs(ogrw).max()
    assert np.max(np.abs(bsp - obsp)) <= 1e-15
    assert np.max(np.abs(sb - osb)) <= 1e-15 * np.abs(osb).max()
    # partition of unity at every Gauss point
    assert np.max(np.abs(bsp.sum(axis=2) - 1.0)) < 1e-14


@pytest.mark.parametrize("ks,ns,z,rmax,h", [(8, 60, 6.0, 20.0, 0.05), (8, 200, 18.0, 40.0, 0.05),
                                            (8, 100, 18.0, 40.0, 0.05), (8, 250, 18.0, 40.0, 0.05),
                                            (9, 150, 54.0, 30.0, 0.05)])
def test_fixed_ns_grid(ks, ns, z, rmax, h):
    from bsr_amd import grid as G
    g = G.fixed_ns_grid(ks, ns, z, rmax, h)
    assert g.ns == ns and len(g.t) == ns + ks
    assert abs(g.t[-1] - rmax) < 1e-12 * rmax
    assert np.all(np.diff(g.t) >= 0)
    # the geometric region really is geometric (the scaling law relies on it)
    zt = g.t * z
    r = zt[ks + g.ml: ks + g.ml + g.me] / zt[ks + g.ml - 1: ks + g.ml + g.me - 1]
    assert np.allclose(r, 1.0 + g.h, rtol=1e-13)


def test_workload_is_deterministic_and_sorted():
    from bsr_amd import synth
    a = synth.c1()
    b = synth.c1()
    assert a.n_records() == b.n_records() > 1000
    for key, la in a.rk_lists.items():
        lb = b.rk_lists[key]
        for f in ("k1", "k2", "k3", "k4", "c"):
            assert np.array_equal(getattr(la, f), getattr(lb, f))
        keys = np.stack([la.k1, la.k2, la.k3, la.k4], 1).astype(np.int64)
        assert len(np.unique(keys, axis=0)) == len(keys)
        order = np.lexsort((keys[:, 3], keys[:, 2], keys[:, 1], keys[:, 0]))
        assert np.array_equal(order, np.arange(len(keys)))
    assert a.nhm == 1200 and a.n_elements == 1200 * 1201 // 2


def test_record_list_merges_duplicates():
    from bsr_amd.synth import RecordList
    r = RecordList.build([3, 1, 3], [1, 1, 1], [2, 2, 2], [0, 0, 0], [0.5, 1.0, 0.25])
    assert list(r.k1) == [1, 3] and np.allclose(r.c, [1.0, 0.75])


def test_bsr_mat_file_round_trip(tmp_path):
    from bsr_amd import synth
    w = synth.make_wave(5, 8, 30, 4.0, 10.0, nch=3, ncp=2, mk=1, nb=2, ntarg=2, fill=0.5, name="rt")
    S, H = O.run_wave(w)
    p = tmp_path / "bsr_mat.001"
    O.write_bsr_mat(p, w, S, H)
    f = O.read_bsr_mat(p)
    assert (f["ns"], f["nch"], f["ncp"], f["mk"], f["ns2"]) == (w.ns, w.nch, w.ncp, w.mk, w.ns)
    for i in range(w.nch):
        assert np.array_equal(f["H"]["diag"][i], H.hcc[:, :, (i + 1) * (i + 2) // 2 - 1])
    assert np.array_equal(f["acf"], H.acf)
    assert f["ec"] == w.ec and np.array_equal(f["etarg"], w.etarg)


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode != 0 and "one GPU" in r.stderr
