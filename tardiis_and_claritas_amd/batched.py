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

    def run_prepared(self, budgets, seed, history_offsets=None, seeds=None):
        """Trace budgets[i] histories of prepared configuration i, all in one
        launch; configuration i uses Philox key seeds[i] (default seed + i)."""
        n = len(budgets)
        if n > self.n_prepared:
            raise ValueError('more budgets than prepared configurations')
        items = (abi.BatchItem * n)()
        for i, b in enumerate(budgets):
            items[i].config = i
            items[i].n_histories = int(b)
            items[i].seed = (int(seeds[i]) if seeds is not None else int(seed) + i) & 0xFFFFFFFFFFFFFFFF
            items[i].history_offset = int(history_offsets[i]) if history_offsets is not None else 0
        out = (abi.Tallies * n)()
        lp = None
        if self.launch:
            l = abi.Launch()
            l.block, l.blocks_per_cu, l.chunk, l.queue_order = (
                int(self.launch.get(k, 0)) for k in ('block', 'blocks_per_cu', 'chunk', 'queue_order'))
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


class Autotuner:
    """Launch-shape autotuner for the persistent transport kernel.

    Candidates are (threads per workgroup, resident workgroups per CU,
    histories per work item, queue order; chunk 0 = item sizes guided by the
    queued work, order 1 = most expensive configuration first).  Each is timed on the
    prepared batch with the real budgets scaled down to a sample large enough
    that the launch is not dominated by its fixed costs and end-of-launch tail
    (the same configurations and the same mix, so divergence is
    representative), and the fastest shape is kept per workload key.  No
    reference counterpart."""

    def __init__(self, candidates=None):
        self.candidates = candidates or [
            dict(block=256, blocks_per_cu=0, chunk=0, queue_order=0),   # library default
            dict(block=256, blocks_per_cu=0, chunk=0, queue_order=1),
            dict(block=192, blocks_per_cu=4, chunk=0, queue_order=1),
            dict(block=128, blocks_per_cu=6, chunk=0, queue_order=1),
            dict(block=256, blocks_per_cu=0, chunk=4096, queue_order=1),
            dict(block=256, blocks_per_cu=0, chunk=16384, queue_order=0),
        ]
        self.cache = {}
        self.last = []

    def tune(self, runner: BatchedRunner, budgets, seed, key=None, sample_fraction=0.25, min_sample=100_000_000):
        """Return the best launch dict for `runner`'s prepared batch."""
        if key is not None and key in self.cache:
            return self.cache[key]
        total = float(sum(budgets))
        f = min(1.0, max(sample_fraction, min_sample / max(total, 1.0)))
        sample = [max(1, int(round(b * f))) for b in budgets]
        saved = runner.launch
        results = []
        for cand in self.candidates:
            runner.launch = cand
            runner.run_prepared(sample, seed)            # warm-up: graph capture for this shape
            runner.run_prepared(sample, seed + 1)
            results.append((sum(sample) / (runner.last_ms * 1e-3), cand))
        runner.launch = saved
        self.last = results
        best = max(results, key=lambda r: r[0])[1]
        if key is not None:
            self.cache[key] = best
        return best


class Pipeline:
    """A stream of batches over the same prepared configurations with `depth`
    launches in flight on one GPU.

    Each stage owns a library context (its own HIP stream, slots and captured
    graph) and a host thread; batch k runs on stage k mod depth.  The device
    overlaps the launches: the next batch's workgroups take the compute units
    that the previous launch frees while its last histories drain, so the
    end-of-launch tail of one batch is filled by the next.  Results are
    returned in batch order.  No reference counterpart."""

    def __init__(self, specs, depth=2, device=0, launch=None):
        if depth < 1:
            raise ValueError('depth must be >= 1')
        self.stages = [BatchedRunner(device, launch) for _ in range(depth)]
        for r in self.stages:
            r.prepare(specs)
        self.last_ms = []

    @property
    def depth(self):
        return len(self.stages)

    def run(self, budgets_list, seeds):
        """budgets_list[k], seeds[k] describe batch k; returns per batch the
        list of tally dicts; `last_ms` holds each launch's device time."""
        from concurrent.futures import ThreadPoolExecutor
        n = len(budgets_list)
        if len(seeds) != n:
            raise ValueError('one seed per batch')
        out, ms = [None] * n, [0.0] * n

        def stage(i):
            r = self.stages[i]
            for k in range(i, n, self.depth):
                out[k] = r.run_prepared(budgets_list[k], seeds[k])
                ms[k] = r.last_ms

        with ThreadPoolExecutor(self.depth) as ex:
            for f in [ex.submit(stage, i) for i in range(min(self.depth, n))]:
                f.result()
        self.last_ms = ms
        return out


def run_config3(n=256, seed=20261203, tune=True, log=print):
    """Config 3: n sampled configurations with log-uniform budgets in one
    persistent batched launch; returns a summary dict."""
    import time
    from . import workloads
    p = workloads.config3_params(n, seed)
    runner = BatchedRunner(0)
    t0 = time.time()
    specs = [workloads.config3_item(p, i) for i in range(n)]
    runner.prepare(specs)
    prep_s = time.time() - t0
    budgets = [int(b) for b in p['budget']]
    out = dict(n_configs=n, histories=int(sum(budgets)), prepare_seconds=prep_s)
    runner.launch = None
    runner.run_prepared(budgets, seed)
    T = runner.run_prepared(budgets, seed + 1)
    out['default_ms'] = runner.last_ms
    out['default_histories_per_s'] = sum(budgets) / (runner.last_ms * 1e-3)
    if tune:
        tuner = Autotuner()
        best = tuner.tune(runner, budgets, seed + 2)
        out['autotune'] = [dict(shape=c, sample_histories_per_s=r) for r, c in tuner.last]
        runner.launch = best
        runner.run_prepared(budgets, seed)
        T = runner.run_prepared(budgets, seed + 1)
        out['tuned_shape'] = best
        out['tuned_ms'] = runner.last_ms
        out['tuned_histories_per_s'] = sum(budgets) / (runner.last_ms * 1e-3)
    out['failures'] = int(sum(int(t['failures'].sum()) for t in T))
    out['status_exit_fraction'] = float(sum(int(t['status'][1]) for t in T) / sum(budgets))
    log(out)
    return out


if __name__ == '__main__':
    import json
    print(json.dumps(run_config3(log=lambda s: None)))
