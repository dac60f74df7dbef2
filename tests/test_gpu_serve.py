"""Serving loop on the GPU: a job's result is the same alone and in a mix,
admission respects the HBM budget, and a failing job does not poison its
slot or the others."""
import numpy as np
import pytest

from tests.conftest import HAS_GPU

pytestmark = pytest.mark.gpu


def _waves(n, nch=16, seed0=500):
    from bsr_amd import synth
    return [synth.make_wave(seed0 + j, 8, 100, 18.0, 40.0, nch=nch + 3 * j, ncp=20 + j, mk=6, nb=5, ntarg=5,
                            fill=0.05, name=f"s{j}", target_seed=9) for j in range(n)]


@pytest.fixture(scope="module")
def server():
    if not HAS_GPU:
        pytest.skip("no HIP device")
    from bsr_amd.serve import Server
    s = Server(nslots=3)
    yield s
    s.close()


def test_alone_equals_in_mix(server):
    ws = _waves(6)
    x = ws[0]
    alone = server.submit(x, keep=True)
    alone.event.wait()
    assert alone.error is None
    jobs = [server.submit(w, keep=True) for w in ws[1:4]]
    mixed = server.submit(x, keep=True)
    jobs += [server.submit(w, keep=True) for w in ws[4:]]
    for j in jobs + [mixed]:
        j.event.wait()
        assert j.error is None, j.error
    # the jobs overlapped on the device
    spans = sorted((j.start, j.done) for j in jobs + [mixed])
    assert any(b[0] < a[1] for a, b in zip(spans, spans[1:]))
    for which in (0, 1):
        a, b = alone.result[which], mixed.result[which]
        assert np.array_equal(a.hcc, b.hcc) and np.array_equal(a.hcb, b.hcb)
        assert np.array_equal(a.hbb, b.hbb) and np.array_equal(a.acf, b.acf)


def test_dirac_alone_equals_in_mix(server):
    """A Dirac wave served alone and in a mix with BS waves on the same
    slots gives identical bits."""
    from bsr_amd import synth_dirac as SD
    d = SD.small(13, npert=2)
    alone = server.submit(d, keep=True)
    alone.event.wait()
    assert alone.error is None, alone.error
    ws = _waves(3, seed0=700)
    jobs = [server.submit(w) for w in ws[:2]]
    mixed = server.submit(d, keep=True)
    jobs.append(server.submit(ws[2]))
    for j in jobs + [mixed]:
        j.event.wait()
        assert j.error is None, j.error
    for key in ("hch", "hcp", "hp", "acf", "htarg"):
        assert np.array_equal(alone.result[key], mixed.result[key]), key
    assert np.abs(alone.result["hch"]).max() > 0


def test_admission_by_footprint():
    if not HAS_GPU:
        pytest.skip("no HIP device")
    from bsr_amd.serve import Server
    ws = _waves(4)
    budget = max(w.footprint_bytes() for w in ws) * 1.5   # room for one job at a time
    s = Server(nslots=3, hbm_budget=budget)
    try:
        jobs = [s.submit(w) for w in ws]
        for j in jobs:
            j.event.wait()
            assert j.error is None
        # at every job start, the footprints of the jobs then running fit the budget
        for j in jobs:
            inflight = sum(k.footprint for k in jobs if k.start <= j.start < k.done)
            assert inflight <= budget, "admitted beyond the HBM budget"
        # and the largest job never ran next to another one (two of them don't fit)
        big = max(jobs, key=lambda k: k.footprint)
        assert all(not (k.start < big.done and big.start < k.done) for k in jobs
                   if k is not big and k.footprint + big.footprint > budget)
    finally:
        s.close()


def test_failed_job_is_isolated(server):
    from bsr_amd import synth
    ws = _waves(3, seed0=700)
    bad = _waves(1, seed0=800)[0]
    lst = bad.rk_lists[(8, 0)]
    lst.k1 = lst.k1.copy()
    lst.k1[0] = (bad.nch + 5) * synth.IBI + 1
    jobs = [server.submit(ws[0]), server.submit(bad), server.submit(ws[1]), server.submit(ws[2])]
    for j in jobs:
        j.event.wait()
    assert jobs[1].error and "channel index out of range" in jobs[1].error
    assert all(j.error is None for j in (jobs[0], jobs[2], jobs[3]))
