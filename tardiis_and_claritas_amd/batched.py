"""Batched Monte Carlo: many configurations in one persistent launch.

Every configuration is prepared into its own device slot; `run` builds one
device work queue of (configuration, history range) items over all of them
and replays a captured hipGraph of the persistent transport kernel
(clr_run_batch).  No reference counterpart: the reference runs one
configuration per simulate() call (CLARITAS.py:7160).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from ._lib import Context
from .model import ClaritasForwardModel


class BatchedRunner:
    def __init__(self, device: int = 0, launch: dict | None = None):
        self.ctx = Context(device)
        self.launch = launch
        self.n_prepared = 0

    def prepare(self, specs):
        """specs: list of (constructor kwargs, Material, concentration g/L)."""
        models = []
        for i, (kw, mat, conc) in enumerate(specs):
            m = ClaritasForwardModel(**kw)
            m._ctx = self.ctx
            m.prepare(mat, conc, slot=i)
            models.append(m)
        self.n_prepared = len(specs)
        return models

    def run_prepared(self, budgets, seed, history_offsets=None):
        n = len(budgets)
        if n > self.n_prepared:
            raise ValueError('more budgets than prepared configurations')
        items = (abi.BatchItem * n)()
        for i, b in enumerate(budgets):
            items[i].config = i
            items[i].n_histories = int(b)
            items[i].seed = (int(seed) + i) & 0xFFFFFFFFFFFFFFFF
            items[i].history_offset = int(history_offsets[i]) if history_offsets is not None else 0
        out = (abi.Tallies * n)()
        lp = None
        if self.launch:
            l = abi.Launch()
            l.block, l.blocks_per_cu, l.chunk = (int(self.launch.get(k, 0)) for k in ('block', 'blocks_per_cu', 'chunk'))
            lp = C.byref(l)
        self.ctx.check(self.ctx.lib.clr_run_batch(self.ctx.h, items, n, lp, out), 'clr_run_batch')
        self.last_ms = float(self.ctx.lib.clr_last_kernel_ms(self.ctx.h))
        return [abi.tallies_to_numpy(out[i]) for i in range(n)]

    def run(self, specs, budgets, seed, history_offsets=None):
        self.prepare(specs)
        return self.run_prepared(budgets, seed, history_offsets)


def flops_per_history_model(counters) -> float:
    """Algorithmic FP32 work of a run under the frozen per-event model
    F = 16 + 30 N_flight + 16 N_null + 60 N_scat + 40 N_iface + 10 N_abs + 45 N_exit
    (FMA = 2, transcendental/sqrt = 1, integer RNG and compares = 0)."""
    c = counters
    ci = abi.C_IDX
    return float(16 * c[ci['histories']] + 30 * c[ci['n_flight']] + 16 * c[ci['n_null']] + 60 * c[ci['n_scat']]
                 + 40 * c[ci['n_iface']] + 10 * c[ci['n_abs']] + 45 * c[ci['n_exit']])


def total_counters(tallies) -> np.ndarray:
    return np.sum([t['counters'] for t in tallies], axis=0)
