"""Continuously batched serving of partial-wave assembly jobs (BSR waves,
and Dirac waves of DBSR_MAT3 through the same slots).

A Server owns `nslots` device contexts.  Each slot has its own HIP stream
and keeps the knot grid, B-spline tables, R^k bands and orbital bases of
the target resident, so a job of the same target only processes its lists
and assembles its S and H.  A dispatcher admits queued jobs in arrival
order as soon as a slot is free and the job's HBM footprint fits the
budget; admitted jobs run concurrently on their slots' streams (the C ABI
releases the GIL), so work of jobs that arrive while others run overlaps on
the device.  Results are deterministic per job: every output element is
summed by one thread in a fixed order, whatever else shares the device.
"""
from __future__ import annotations

import threading
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from .core import Context


def is_dirac(wave) -> bool:
    """Dirac (DBSR) partial waves carry two-component orbitals pq."""
    return hasattr(wave, "pq") and hasattr(wave, "ipbs")


@dataclass
class Job:
    wave: object
    arrival: float                      # seconds on the server clock
    keep: bool = False                  # copy S and H back to the host
    footprint: int = 0
    slot: int = -1
    start: float = 0.0
    done: float = 0.0
    device_ms: float = 0.0
    error: str | None = None
    result: tuple | None = None
    event: threading.Event = field(default_factory=threading.Event)

    @property
    def latency(self):
        return self.done - self.arrival


class Server:
    def __init__(self, nslots: int = 4, hbm_budget: float = 200e9, precision: int = 0, nslices: int = 0,
                 device: int = 0, device_lists="auto"):
        # device_lists: True, a BS job's R^k lists go to device memory and are
        # decoded there (bsrg_wave_assemble_resident: the host out of the
        # job's way, lowest latency); False, the host workers of its slot
        # decode them (the device does less per job: highest capacity when
        # jobs queue); "auto", the first when no job waits in the queue as
        # the job starts, else the second.  Both give the same bits.
        if device_lists not in (True, False, "auto"):
            raise ValueError("device_lists must be True, False or 'auto'")
        self.device_lists = device_lists
        self.device_list_jobs = 0
        self.slots = [Context(device) for _ in range(nslots)]
        for c in self.slots:
            c.set_precision(precision, nslices if precision else 0)
        self.budget = hbm_budget
        self.used = 0
        self.queue: deque[Job] = deque()
        self.free = list(range(nslots))
        self.cv = threading.Condition()
        self.t0 = time.perf_counter()
        self.stopping = False
        self.jobs_done: list[Job] = []
        self.workers = []
        self.inbox = [deque() for _ in range(nslots)]
        for s in range(nslots):
            th = threading.Thread(target=self._worker, args=(s,), daemon=True)
            th.start()
            self.workers.append(th)
        self.dispatcher = threading.Thread(target=self._dispatch, daemon=True)
        self.dispatcher.start()

    def now(self):
        return time.perf_counter() - self.t0

    def submit(self, wave, arrival: float | None = None, keep: bool = False) -> Job:
        job = Job(wave, self.now() if arrival is None else arrival, keep)
        job.footprint = int(wave.footprint_bytes())
        if job.footprint > self.budget:
            raise ValueError(f"job needs {job.footprint} bytes, more than the HBM budget {self.budget:.0f}")
        with self.cv:
            if self.stopping:
                raise RuntimeError("server is closing")
            self.queue.append(job)
            self.cv.notify_all()
        return job

    def _dispatch(self):
        while True:
            with self.cv:
                while not (self.queue and self.free and self.used + self.queue[0].footprint <= self.budget) \
                        and not (self.stopping and not self.queue):
                    self.cv.wait()
                if self.stopping and not self.queue:
                    return
                if not (self.queue and self.free and self.used + self.queue[0].footprint <= self.budget):
                    continue
                job = self.queue.popleft()
                s = self.free.pop(0)
                self.used += job.footprint
                job.slot = s
                self.inbox[s].append(job)
                self.cv.notify_all()

    def _worker(self, s: int):
        ctx = self.slots[s]
        while True:
            with self.cv:
                # on close, a worker leaves only once no job is queued: the
                # dispatcher may still hand queued jobs to this slot
                while not self.inbox[s] and not (self.stopping and not self.queue):
                    self.cv.wait()
                if not self.inbox[s]:
                    return
                job = self.inbox[s].popleft()
                dev_lists = self.device_lists is True or (self.device_lists == "auto" and not self.queue)
                self.device_list_jobs += int(dev_lists)
            job.start = self.now()
            try:
                if is_dirac(job.wave):
                    # a Dirac (DBSR) wave: grid, orbitals and lists of the four families
                    ctx.dirac_wave(job.wave)
                    job.device_ms = float(ctx.dirac_timings()[4])
                    if job.keep:
                        job.result = ctx.dirac_get()
                else:
                    if dev_lists:
                        ctx.lists_upload(job.wave)
                        ctx.wave_assemble_resident(job.wave)
                    else:
                        ctx.wave_assemble(job.wave)
                    job.device_ms = ctx.wave_timings()["total"]
                    if job.keep:
                        job.result = (ctx.wave_get(0), ctx.wave_get(1))
            except Exception as e:  # a failed job does not poison the slot
                job.error = str(e)
            job.done = self.now()
            with self.cv:
                self.used -= job.footprint
                self.free.append(s)
                self.jobs_done.append(job)
                self.cv.notify_all()
            job.event.set()

    def run_stream(self, stream, keep: bool = False):
        """Submit [(arrival_s, wave)] at their arrival times; wait for all."""
        jobs = []
        base = self.now()
        for arr, w in stream:
            delay = base + arr - self.now()
            if delay > 0:
                time.sleep(delay)
            jobs.append(self.submit(w, arrival=base + arr, keep=keep))
        for j in jobs:
            j.event.wait()
        return jobs

    def close(self):
        """Stop accepting jobs, run every queued job to completion, then
        release the slots."""
        with self.cv:
            self.stopping = True
            self.cv.notify_all()
        for th in self.workers:
            th.join()
        self.dispatcher.join()
        for c in self.slots:
            c.close()


def summarize(jobs) -> dict:
    lat = np.array([j.latency for j in jobs]) * 1e3
    span = max(j.done for j in jobs) - min(j.arrival for j in jobs)
    elems = sum(j.wave.n_elements for j in jobs)
    service = np.array([j.done - j.start for j in jobs]) * 1e3
    return dict(jobs=len(jobs), p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                elements_per_s=elems / span, span_s=span, failed=sum(j.error is not None for j in jobs),
                service_ms_mean=float(service.mean()), device_ms_mean=float(np.mean([j.device_ms for j in jobs])))
