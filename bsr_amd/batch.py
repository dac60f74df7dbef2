"""Batches of uneven partial waves run as one hipGraph.

Each wave gets its own device context (stream, grid tables, R^k bands,
bases, matrices).  `prepare` does the host list processing and leaves every
input resident; the device work of all waves is then captured once into a
single hipGraph (one branch per wave) and every `run` replays the whole
batch with one graph launch.  The GEMM engine of each wave's exchange
contraction is chosen by the autotuner among the hand-written kernels (the
capturable ones) for that wave's shape.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as L
from .core import Context


class Batch:
    def __init__(self, waves, precision: int = 0, nslices: int = 0, device: int = 0):
        self.waves = list(waves)
        self.ctx = [Context(device) for _ in self.waves]
        for c in self.ctx:
            c.set_capture_safe(True)
            c.set_precision(precision, nslices if precision else 0)
        self._graph = C.c_void_p()

    def prepare(self):
        """Host lists and uploads, then one eager pass per wave (device time
        of each wave's launch step alone, and the host scalars)."""
        for c, w in zip(self.ctx, self.waves):
            c.wave_prepare(w)
        self.eager_ms = []
        for c in self.ctx:
            c.wave_launch()
            c.wave_finish()
            t = c.wave_timings()
            self.eager_ms.append(t["gemm"] + t["terms"])

    def capture(self):
        arr = (C.c_void_p * len(self.ctx))(*[c._p for c in self.ctx])
        L.check(L.lib().bsrg_graph_capture(len(self.ctx), arr, C.byref(self._graph)))

    def run(self):
        """One replay of the whole batch; returns (per-wave ms, graph ms)."""
        L.check(L.lib().bsrg_graph_launch(self._graph))
        per = np.zeros(len(self.ctx), order="F")
        tot = C.c_double()
        L.check(L.lib().bsrg_graph_sync(self._graph, per, C.byref(tot)))
        return per, tot.value

    def engines(self):
        return [c.wave_engine() for c in self.ctx]

    def n_elements(self):
        return sum(w.n_elements for w in self.waves)

    def close(self):
        if self._graph:
            L.lib().bsrg_graph_destroy(self._graph)
            self._graph = C.c_void_p()
        for c in self.ctx:
            c.close()
