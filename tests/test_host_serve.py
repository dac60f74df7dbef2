"""Host logic of the serving loop (bsr_amd/serve.py) with stand-in slot
contexts: admission by footprint, closing with jobs still queued, failure
isolation.  No device is used."""
import threading
import time

import pytest

from bsr_amd import serve


class _FakeCtx:
    """Stands in for a device context: wave_assemble (and its form with the
    lists in device memory) sleeps for the wave's duration (or raises),
    wave_timings reports it."""
    active = 0
    peak = 0
    lock = threading.Lock()

    def __init__(self, device=0):
        self.last = 0.0

    def set_precision(self, mode, nslices):
        pass

    def wave_assemble(self, w):
        with _FakeCtx.lock:
            _FakeCtx.active += 1
            _FakeCtx.peak = max(_FakeCtx.peak, _FakeCtx.active)
        try:
            if w.fail:
                raise RuntimeError("bad wave")
            time.sleep(w.dur)
            self.last = w.dur
        finally:
            with _FakeCtx.lock:
                _FakeCtx.active -= 1

    def lists_upload(self, w):
        pass

    def wave_assemble_resident(self, w):
        self.wave_assemble(w)

    def wave_timings(self):
        return {"total": self.last * 1e3}

    def wave_get(self, which):
        return which

    def close(self):
        pass


class _Wave:
    def __init__(self, dur=0.01, fp=1, fail=False):
        self.dur, self.fp, self.fail, self.n_elements = dur, fp, fail, 10

    def footprint_bytes(self):
        return self.fp


@pytest.fixture
def fake(monkeypatch):
    monkeypatch.setattr(serve, "Context", _FakeCtx)
    _FakeCtx.active = _FakeCtx.peak = 0
    return _FakeCtx


def test_close_runs_queued_jobs(fake):
    s = serve.Server(nslots=2, hbm_budget=100)
    jobs = [s.submit(_Wave(0.02)) for _ in range(9)]
    s.close()   # most jobs are still queued here
    assert all(j.event.is_set() for j in jobs)
    assert all(j.error is None for j in jobs)
    with pytest.raises(RuntimeError):
        s.submit(_Wave())


def test_admission_by_footprint(fake):
    s = serve.Server(nslots=4, hbm_budget=10)
    jobs = [s.submit(_Wave(0.02, fp=6)) for _ in range(4)]   # two never fit together
    for j in jobs:
        assert j.event.wait(10)
    s.close()
    assert fake.peak == 1
    with pytest.raises(ValueError):
        serve.Server.submit(s, _Wave(fp=11))


def test_failed_job_leaves_slot_usable(fake):
    s = serve.Server(nslots=1, hbm_budget=100)
    bad = s.submit(_Wave(fail=True))
    good = s.submit(_Wave(0.01))
    assert good.event.wait(10) and bad.event.wait(10)
    s.close()
    assert bad.error == "bad wave" and good.error is None
    st = serve.summarize([good])
    assert st["failed"] == 0 and st["jobs"] == 1
