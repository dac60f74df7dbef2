"""Batches of uneven partial waves run as one hipGraph.

Each wave gets its own device context (stream, grid tables, R^k bands,
bases, matrices).  `prepare` does the host list processing and leaves every
input resident; the device work of all waves is then captured once into a
single hipGraph and every `run` replays the whole batch with one graph
launch.  With `persistent` (the default) the exchange products of all waves
run in one persistent launch that takes work items of every wave from one
counter, its granularity chosen by the library's autotuner on the first
capture of the batch's shape; otherwise each wave's launch step is one
branch of the graph.  Waves whose product is not the sparse FP64 one keep
their own branch, their exchange engine chosen by the per-shape autotuner
among the capturable kernels.
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

    def capture(self, persistent: bool = True):
        if self._graph:
            L.lib().bsrg_graph_destroy(self._graph)
            self._graph = C.c_void_p()
        arr = (C.c_void_p * len(self.ctx))(*[c._p for c in self.ctx])
        f = L.lib().bsrg_graph_capture_batched if persistent else L.lib().bsrg_graph_capture
        L.check(f(len(self.ctx), arr, C.byref(self._graph)))

    def batch_info(self):
        """(waves in the persistent launch, blocks per work item chosen by
        the autotuner; 0 = the waves' own grids)."""
        nb, bpi = C.c_int(), C.c_int()
        L.check(L.lib().bsrg_graph_batch_info(self._graph, C.byref(nb), C.byref(bpi)))
        return nb.value, bpi.value

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
