"""Learned surrogate of the detector signals (config 4; no reference counterpart).

An ensemble of MLPs (default 5 members, 8 -> 256 -> 256 -> 256 -> 3, GELU) maps
the eight configuration parameters

    log10 concentration, PSD median (um), sigma_g, m_r, log10 m_i,
    wavelength (nm), vortex depth (mm), Rouse number

to the 90 deg, 0 deg and 135 deg detector signals (symmetrised exit-position
window counts per launched history).  It is trained on the GPU by the library's own
BF16-MFMA forward/backward kernels (clr_mlp_*), on batched Monte Carlo of the
config-3 parameter space.  `Surrogate.query` answers from the network where the
members agree and falls back to Monte Carlo where their spread exceeds a
threshold.

The NTU-equivalent reading is K * S90 / S0 (ratio nephelometry); its relative
error is the error of S90/S0.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import time
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import abi, workloads
from ._lib import Context

FEATURES = ('log10_conc_mg_l', 'log10_median_um', 'sigma_g', 'm_r', 'log10_m_i', 'wavelength_nm', 'vortex_depth_mm',
            'rouse')
SIGNALS = ('s90', 's0', 's135')
P = C.POINTER


def _fp(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class MlpEnsemble:
    """Device ensemble behind clr_mlp_* (weights live in the HIP context)."""

    def __init__(self, ctx: Context, n_members=5, in_dim=8, hidden=256, out_dim=3, n_hidden_layers=3, seed=0):
        self.ctx = ctx
        self.shape = abi.MlpShape(n_members, in_dim, hidden, out_dim, n_hidden_layers, 0)
        ctx.check(ctx.lib.clr_mlp_init(ctx.h, C.byref(self.shape), C.c_uint64(seed)), 'clr_mlp_init')
        self.n_params = int(ctx.lib.clr_mlp_param_count(ctx.h))

    @property
    def n_members(self):
        return self.shape.n_members

    def get_params(self):
        out = np.empty(self.n_params, dtype=np.float32)
        self.ctx.check(self.ctx.lib.clr_mlp_get_params(self.ctx.h, out.ctypes.data_as(P(C.c_float)), out.size),
                       'clr_mlp_get_params')
        return out

    def set_params(self, flat):
        flat = _fp(flat)
        self.ctx.check(self.ctx.lib.clr_mlp_set_params(self.ctx.h, flat.ctypes.data_as(P(C.c_float)), flat.size),
                       'clr_mlp_set_params')

    @staticmethod
    def _weights(w, y):
        if w is None:
            return None, None
        w = _fp(w)
        if w.shape != y.shape:
            raise ValueError(f'weights must have the shape of the targets {y.shape}, got {w.shape}')
        return w, w.ctypes.data_as(P(C.c_float))

    def grad(self, x, y, w=None):
        """Per-member losses and the flat gradient of the (weighted) MSE on one batch."""
        x, y = _fp(x), _fp(y)
        w, wp = self._weights(w, y)
        loss = np.empty(self.n_members, dtype=np.float32)
        g = np.empty(self.n_params, dtype=np.float32)
        self.ctx.check(self.ctx.lib.clr_mlp_grad(self.ctx.h, x.ctypes.data_as(P(C.c_float)),
                                                 y.ctypes.data_as(P(C.c_float)), wp, len(x),
                                                 loss.ctypes.data_as(P(C.c_float)), g.ctypes.data_as(P(C.c_float))),
                       'clr_mlp_grad')
        return loss, g

    def train(self, x, y, batch=512, epochs=10, lr=1e-3, seed=0, w=None):
        """Adam epochs over (x, y); `w` holds per-element loss weights."""
        x, y = _fp(x), _fp(y)
        w, wp = self._weights(w, y)
        loss = np.empty(self.n_members, dtype=np.float32)
        self.ctx.check(self.ctx.lib.clr_mlp_train(self.ctx.h, x.ctypes.data_as(P(C.c_float)),
                                                  y.ctypes.data_as(P(C.c_float)), wp, len(x), int(batch),
                                                  int(epochs), C.c_float(lr), C.c_uint64(seed),
                                                  loss.ctypes.data_as(P(C.c_float))),
                       'clr_mlp_train')
        return loss

    def predict(self, x):
        x = _fp(x)
        mean = np.empty((len(x), self.shape.out_dim), dtype=np.float32)
        std = np.empty_like(mean)
        self.ctx.check(self.ctx.lib.clr_mlp_infer(self.ctx.h, x.ctypes.data_as(P(C.c_float)), len(x),
                                                  mean.ctypes.data_as(P(C.c_float)), std.ctypes.data_as(P(C.c_float))),
                       'clr_mlp_infer')
        return mean, std

    def time_infer(self, n=65536, repeats=20):
        ms = C.c_double()
        self.ctx.check(self.ctx.lib.clr_mlp_time_infer(self.ctx.h, int(n), int(repeats), C.byref(ms)),
                       'clr_mlp_time_infer')
        return ms.value


# ---------------------------------------------------------------- features --
def features(p) -> np.ndarray:
    """[n][8] raw features from a config-3 style parameter dict."""
    return np.stack([np.log10(p['conc_mg_l']), np.log10(p['median_um']), p['sigma_g'], p['m_r'], np.log10(p['m_i']),
                     p['wavelength_nm'], p['vortex_depth_mm'], p['rouse']], axis=1).astype(np.float64)


def signals_from_tallies(T, n_histories):
    """S90, S0, S135: symmetrised exit-position window counts (the legacy
    +-acceptance rule at 90 and 0 deg, the aux window at 135 deg) per launched
    history.  The collimated hardware bores accept ~1e-7 of the launched
    photons, too few for Monte Carlo labels at affordable photon counts."""
    det = T['det'].sum(0)           # [4][18]: hw native, hw mirror, legacy native, legacy mirror
    auxw = T['aux_window'].sum(0)   # [2][n_aux]
    s90 = 0.5 * (det[2, 9] + det[3, 9]) / n_histories
    s0 = 0.5 * (det[2, 0] + det[3, 0]) / n_histories
    s135 = 0.5 * (auxw[0, 0] + auxw[1, 0]) / n_histories
    return np.array([s90, s0, s135], dtype=np.float64)


# ---------------------------------------------------------------- dataset --
def sample_params(n, seed):
    """Configuration parameters from the config-3 distributions."""
    p = workloads.config3_params(n, seed)
    p.pop('budget')
    return p


def monte_carlo(params, photons, seed, batch=256, runner=None, progress=None):
    """Signals [n][3] of every configuration by batched Monte Carlo, `photons`
    histories each (scalar or per-configuration array)."""
    from .batched import BatchedRunner
    n = len(params['median_um'])
    photons = np.broadcast_to(np.asarray(photons, dtype=np.int64), (n,))
    runner = runner or BatchedRunner(0)
    out = np.empty((n, 3))
    for b0 in range(0, n, batch):
        idx = range(b0, min(n, b0 + batch))
        specs = [workloads.config3_item(params, i) for i in idx]
        runner.prepare(specs)
        T = runner.run_prepared([int(photons[i]) for i in idx], seed + b0)
        for k, i in enumerate(idx):
            out[i] = signals_from_tallies(T[k], int(photons[i]))
        if progress:
            progress(min(n, b0 + batch), n)
    return out


OPTICAL = ('median_um', 'sigma_g', 'm_r', 'm_i', 'wavelength_nm')
TRANSPORT = ('conc_mg_l', 'vortex_depth_mm', 'rouse')


def transport_params(n, g):
    """Concentration, vortex depth and Rouse number from the config-3 ranges."""
    return dict(conc_mg_l=np.exp(g.uniform(np.log(10.0), np.log(10000.0), n)),
                vortex_depth_mm=g.uniform(0.0, 6.0, n), rouse=g.uniform(0.05, 1.0, n))


def generate(n_optical, n_variants, photons, seed, runner=None, progress=None):
    """Dataset of n_optical x n_variants configurations from the config-3
    distributions: the five optical parameters are drawn per optical
    configuration and the three transport parameters afresh for every variant.
    A batch holds up to 256 optical configurations, one per device slot, and
    the variants re-prepare the same slots, so each slot's Mie tables are
    computed once.  Returns (params dict, signals [n][3])."""
    from .batched import BatchedRunner
    runner = runner or BatchedRunner(0)
    p_opt = workloads.config3_params(n_optical, seed)
    g = np.random.Generator(np.random.PCG64(seed + 7919))
    cols = {k: [] for k in OPTICAL + TRANSPORT}
    sig = []
    done, total = 0, n_optical * n_variants
    for a0 in range(0, n_optical, 256):
        sel = np.arange(a0, min(n_optical, a0 + 256))
        for v in range(n_variants):
            params = {k: np.asarray(p_opt[k])[sel] for k in OPTICAL}
            params.update(transport_params(len(sel), g))
            sig.append(monte_carlo(params, photons, seed + 1_000_003 * (v + 1) + a0, batch=256, runner=runner))
            for k in cols:
                cols[k].append(params[k])
            done += len(sel)
            if progress:
                progress(done, total)
    return {k: np.concatenate(v) for k, v in cols.items()}, np.concatenate(sig)


# ---------------------------------------------------------------- surrogate --
@dataclass
class Normaliser:
    x_mean: np.ndarray
    x_std: np.ndarray
    y_mean: np.ndarray
    y_std: np.ndarray

    def to_dict(self):
        return {k: v.tolist() for k, v in self.__dict__.items()}


FLOOR = 1e-9   # signal floor before the log transform
REL_FLOOR = 0.003   # relative error floor added to every label's Monte Carlo error in the weights


def label_variance(S, n_hist):
    """Monte Carlo variance of the three log targets of Surrogate.targets from
    Poisson window counts c = S * n_hist: log(S90/S0) -> 1/c90 + 1/c0,
    log S0 -> 1/c0, log(S135/S0) -> 1/c135 + 1/c0."""
    c = np.maximum(np.asarray(S, dtype=np.float64) * np.asarray(n_hist, dtype=np.float64)[:, None], 1.0)
    return np.stack([1.0 / c[:, 0] + 1.0 / c[:, 1], 1.0 / c[:, 1], 1.0 / c[:, 2] + 1.0 / c[:, 1]], axis=1)


def label_weights(S, n_hist, rel_floor=REL_FLOOR):
    """Inverse-variance loss weights [n][3], each column scaled to mean 1.  The
    floor keeps labels with very many counts from dominating once the
    network's own error exceeds their statistical error."""
    w = 1.0 / (label_variance(S, n_hist) + rel_floor ** 2)
    return (w / w.mean(0)).astype(np.float32)


class Surrogate:
    """Feature/target transforms around an MlpEnsemble, with MC fallback."""

    def __init__(self, ctx: Context, n_members=5, hidden=256, n_hidden_layers=3, seed=0):
        self.ens = MlpEnsemble(ctx, n_members, len(FEATURES), hidden, len(SIGNALS), n_hidden_layers, seed)
        self.norm: Normaliser | None = None

    def _x(self, F):
        return (F - self.norm.x_mean) / self.norm.x_std

    @staticmethod
    def targets(S):
        """log(S90/S0), log S0, log(S135/S0): the first is the NTU reading."""
        S = np.maximum(S, FLOOR)
        return np.stack([np.log(S[:, 0] / S[:, 1]), np.log(S[:, 1]), np.log(S[:, 2] / S[:, 1])], axis=1)

    @staticmethod
    def signals(Y):
        s0 = np.exp(Y[:, 1])
        return np.stack([np.exp(Y[:, 0]) * s0, s0, np.exp(Y[:, 2]) * s0], axis=1)

    def fit(self, F, S, epochs=40, batch=512, lr=1e-3, seed=0, schedule=(1.0, 0.3, 0.1), n_hist=None,
            rel_floor=REL_FLOOR):
        """Train on raw features F [n][8] and signals S [n][3] (log-ratio
        targets).  With the histories behind each label (`n_hist`), every
        target is weighted by its inverse Monte Carlo variance (see
        label_weights).  The learning rate steps down over `schedule` (equal
        shares of the epochs)."""
        Y = self.targets(S)
        self.norm = Normaliser(F.mean(0), F.std(0) + 1e-12, Y.mean(0), Y.std(0) + 1e-12)
        x, y = self._x(F), (Y - self.norm.y_mean) / self.norm.y_std
        w = None if n_hist is None else label_weights(S, n_hist, rel_floor)
        losses = []
        per = max(1, epochs // len(schedule))
        for k, f in enumerate(schedule):
            losses.append(self.ens.train(x, y, batch, per, lr * f, seed + k, w=w).tolist())
        return losses

    def save(self, path):
        """Write the ensemble parameters (FP32 master copy) and the
        feature/target normaliser to an .npz file."""
        np.savez(path, params=self.ens.get_params(), **{k: v for k, v in self.norm.__dict__.items()})

    @classmethod
    def load(cls, ctx: Context, path, n_members=5, hidden=256, n_hidden_layers=3):
        """A trained surrogate from a file written by `save` (or run_pipeline's
        weights.npz)."""
        d = np.load(path)
        sur = cls(ctx, n_members=n_members, hidden=hidden, n_hidden_layers=n_hidden_layers)
        sur.ens.set_params(np.asarray(d['params'], dtype=np.float32))
        sur.norm = Normaliser(*(np.asarray(d[k], dtype=np.float64) for k in ('x_mean', 'x_std', 'y_mean', 'y_std')))
        return sur

    def predict(self, F):
        """Signals [n][3] and the members' spread of the log targets [n][3];
        column 0 is the relative spread of the NTU reading S90/S0."""
        m, s = self.ens.predict(self._x(F))
        return self.signals(m * self.norm.y_std + self.norm.y_mean), s * self.norm.y_std

    def query(self, params, threshold=None, mc_photons=None, rel=0.01, seed=1, runner=None):
        """Signals for a parameter dict: network where the ensemble spread of
        log(S90/S0) is within `threshold`, batched Monte Carlo elsewhere.
        `threshold` defaults to the committed fallback threshold
        (surrogate_eval.json, bench_fallback_threshold); the Monte Carlo
        answers get `mc_photons` histories each, by default the histories
        for a relative error `rel` on S90/S0 sized from the network's own
        prediction (clipped to [2e5, 2e9]).  Returns (signals, used_mc
        mask)."""
        if threshold is None:
            threshold = default_fallback_threshold()
        F = features(params)
        S, spread = self.predict(F)
        mc = spread[:, 0] > threshold
        if np.any(mc):
            sub = {k: np.asarray(v)[mc] for k, v in params.items()}
            n = (np.clip(mc_matched_photons(S[mc], rel), 200_000, 2_000_000_000).astype(np.int64)
                 if mc_photons is None else mc_photons)
            S[mc] = monte_carlo(sub, n, seed, runner=runner)
        return S, mc


def default_fallback_threshold():
    """The committed fallback threshold: the held-out 99th percentile of the
    ensemble spread of log(S90/S0) for the shipped weights."""
    ev = Path(__file__).resolve().parent / 'surrogate_eval.json'
    return float(json.loads(ev.read_text())['bench_fallback_threshold'])


def ntu_relative_error(S_pred, S_ref):
    """|dNTU|/NTU with NTU = K S90/S0."""
    return np.abs((S_pred[:, 0] / S_pred[:, 1]) / (S_ref[:, 0] / S_ref[:, 1]) - 1.0)


# ---------------------------------------------------------------- pipeline --
def photon_budget(conc_mg_l, counts=200.0, p90_per_mg_l=1e-7, lo=200_000, hi=200_000_000):
    """Histories per configuration so that the 90 deg window collects about
    `counts` photons: its single-scattering yield grows ~linearly with the
    concentration (p90 ~ 1e-7 per mg/L on the config-3 space)."""
    n = counts / (p90_per_mg_l * np.asarray(conc_mg_l, dtype=float))
    return np.clip(n, lo, hi).astype(np.int64)


def adaptive_budget(s90_pilot, pilot_photons, counts, lo=200_000, hi=400_000_000, s0_pilot=None):
    """Histories per configuration for a label of about `counts` effective
    counts, from pilot estimates of the window yields (pilot counts /
    `pilot_photons`).  With the 90 deg yield alone the 90 deg window collects
    about `counts` photons.  With the 0 deg yield as well, the relative
    variance of the NTU reading S90/S0 (disjoint windows, so
    1/(n S90) + 1/(n S0)) is brought to 1/counts:
    n = counts (1/S90 + 1/S0); dense suspensions, whose transmitted beam is
    nearly extinct, then get the histories their 0 deg label needs.  The
    yields vary about a hundredfold across the optical parameters at one
    concentration, so a budget from the concentration alone leaves the
    weakly scattering, strongly absorbing configurations with a few hundred
    counts (and the noisiest labels) while spending far more than needed on
    the rest.  A pilot with no count is taken to have half a count."""
    floor = 0.5 / float(pilot_photons)
    inv = 1.0 / np.maximum(np.asarray(s90_pilot, dtype=float), floor)
    if s0_pilot is not None:
        inv = inv + 1.0 / np.maximum(np.asarray(s0_pilot, dtype=float), floor)
    return np.clip(counts * inv, lo, hi).astype(np.int64)


def mc_matched_photons(S, rel=0.01):
    """Histories for a relative standard error `rel` on S90/S0 (binomial
    counts; S0 and S90 are disjoint windows)."""
    s90, s0 = np.maximum(S[:, 0], 1e-12), np.maximum(S[:, 1], 1e-12)
    return (1.0 / s90 + 1.0 / s0) / rel ** 2


def load_dataset(path, counts=None, test_counts=None):
    """(train params, train signals, train histories, test params, test
    signals, test histories) from a dataset.npz written by run_pipeline.
    Files without per-sample histories recompute them from the counts they
    were generated with."""
    d = np.load(path)
    out = []
    for part, cnt in (('train', counts), ('test', test_counts)):
        prm = {k: d[f'{part}_{k}'] for k in OPTICAL + TRANSPORT}
        if f'{part}_hist' in d.files:
            hist = d[f'{part}_hist']
        elif cnt is not None:
            hist = photon_budget(prm['conc_mg_l'], cnt)
        else:
            raise ValueError(f'{path}: no {part}_hist; give the counts it was generated with')
        out += [prm, d[f'{part}_S'], hist]
    return tuple(out)


def load_spot(path):
    """(params, signals, histories) of a spot-check set: the held-out part
    of a dataset.npz, or the spot_* arrays of a file written by
    write_heldout."""
    d = np.load(path)
    if 'spot_S' in d.files:
        return {k: d[f'spot_{k}'] for k in OPTICAL + TRANSPORT}, d['spot_S'], d['spot_hist']
    return load_dataset(path)[3:]


def write_heldout(dataset_path, out_path):
    """Copy the held-out part of a dataset.npz (parameters, signals and the
    histories behind them) to `out_path` as the spot_* arrays the benchmark's
    live surrogate evaluation reads."""
    d = np.load(dataset_path)
    out = {f'spot_{k}': np.asarray(d[f'test_{k}'], dtype=np.float64) for k in OPTICAL + TRANSPORT}
    out['spot_S'] = np.asarray(d['test_S'], dtype=np.float64)
    out['spot_hist'] = np.asarray(d['test_hist'], dtype=np.int64)
    np.savez_compressed(out_path, **out)
    return len(out['spot_S'])


def _cat(parts):
    return {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}


def run_pipeline(out_dir, n_optical=64, n_variants=64, n_test_optical=16, n_test_variants=32, counts=200.0,
                 test_counts=2000.0, epochs=60, seed=20261204, log=print, datasets=(), weighted=True,
                 generate_only=False, hi=200_000_000, spot_datasets=(), adaptive_counts=0.0,
                 pilot_photons=1_000_000, time_budget=None, ratio_budget=False, adaptive_test=False):
    """Generate (n_optical x n_variants training and n_test_optical x
    n_test_variants held-out configurations; 0 skips), merge earlier
    `datasets` (paths of dataset.npz), train, evaluate on the held-out part
    and on the held-out parts of `spot_datasets` (Monte Carlo spot checks at
    high counts, scored separately, with the Monte Carlo fallback run on
    them); writes summary.json, weights.npz and dataset.npz (this run's own
    samples).  With `adaptive_counts` > 0 the training configurations get
    adaptive_budget histories from a pilot of `pilot_photons` each (its own
    random streams, so a label is unbiased given its budget); otherwise the
    concentration-only photon_budget with `counts`; `ratio_budget` sizes the
    adaptive budget for the S90/S0 ratio instead of S90, and `adaptive_test`
    gives the held-out configurations that budget with `test_counts`.  With
    `time_budget`
    (seconds) the training generation stops after the first batch that ends
    past it and keeps the configurations done so far."""
    from .batched import BatchedRunner
    out = Path(out_dir)
    out.mkdir(parents=True, exist_ok=True)
    runner = BatchedRunner(0)
    t0 = time.time()

    def gen(n_opt, n_var, cnt, sd, adapt=0.0, budget_s=None):
        p_opt = workloads.config3_params(n_opt, sd)
        g = np.random.Generator(np.random.PCG64(sd + 7919))
        cols = {k: [] for k in OPTICAL + TRANSPORT}
        sig, hists, hist, mc_s = [], [], 0, 0.0
        for a0 in range(0, n_opt, 256):
            sel = np.arange(a0, min(n_opt, a0 + 256))
            for v in range(n_var):
                if budget_s is not None and time.time() - t0 > budget_s:
                    break
                params = {k: np.asarray(p_opt[k])[sel] for k in OPTICAL}
                params.update(transport_params(len(sel), g))
                specs = [workloads.config3_item(params, i) for i in range(len(sel))]
                runner.prepare(specs)
                if adapt > 0:
                    Tp = runner.run_prepared([pilot_photons] * len(sel), sd + 7_000_003 * (v + 1) + a0)
                    hist += pilot_photons * len(sel)
                    mc_s += runner.last_ms * 1e-3
                    Sp = np.stack([signals_from_tallies(Tp[k], pilot_photons) for k in range(len(sel))])
                    budget = adaptive_budget(Sp[:, 0], pilot_photons, adapt, hi=hi,
                                             s0_pilot=Sp[:, 1] if ratio_budget else None)
                else:
                    budget = photon_budget(params['conc_mg_l'], cnt, hi=hi)
                T = runner.run_prepared(budget.tolist(), sd + 1_000_003 * (v + 1) + a0)
                hist += int(budget.sum())
                mc_s += runner.last_ms * 1e-3
                sig.append(np.stack([signals_from_tallies(T[k], int(budget[k])) for k in range(len(sel))]))
                hists.append(budget)
                for k in cols:
                    cols[k].append(params[k])
                log(f'generate: {sum(len(c) for c in cols["rouse"])}/{n_opt * n_var} configs, '
                    f'{hist:.3g} histories, {time.time() - t0:.0f} s')
        if not sig:
            return {k: np.zeros(0) for k in cols}, np.zeros((0, 3)), np.zeros(0, dtype=np.int64), 0, 0.0
        return ({k: np.concatenate(v) for k, v in cols.items()}, np.concatenate(sig), np.concatenate(hists), hist,
                mc_s)

    P_tr, S_tr, n_tr, h_tr, s_tr = gen(n_optical, n_variants, counts, seed, adaptive_counts, time_budget)
    P_te, S_te, n_te, h_te, s_te = gen(n_test_optical, n_test_variants, test_counts, seed + 1,
                                       test_counts if adaptive_test else 0.0)
    np.savez(out / 'dataset.npz', **{f'train_{k}': v for k, v in P_tr.items()}, train_S=S_tr, train_hist=n_tr,
             **{f'test_{k}': v for k, v in P_te.items()}, test_S=S_te, test_hist=n_te)
    mc_rate = (h_tr + h_te) / (s_tr + s_te) if (s_tr + s_te) > 0 else None
    if generate_only:
        summary = dict(n_train=len(S_tr), n_test=len(S_te), train_histories=int(n_tr.sum()),
                       test_histories=int(n_te.sum()), mc_histories_per_s=mc_rate, seconds=time.time() - t0)
        (out / 'summary.json').write_text(json.dumps(summary, indent=1))
        log(json.dumps(summary))
        return summary
    for path in datasets:
        a_P, a_S, a_n, b_P, b_S, b_n = load_dataset(path, counts, test_counts)
        P_tr, S_tr, n_tr = _cat([P_tr, a_P]) if len(S_tr) else a_P, np.concatenate([S_tr, a_S]), \
            np.concatenate([n_tr, a_n])
        P_te, S_te, n_te = _cat([P_te, b_P]) if len(S_te) else b_P, np.concatenate([S_te, b_S]), \
            np.concatenate([n_te, b_n])
    if mc_rate is None:      # no generation in this run: time a held-out batch for the MC cost
        sel = {k: v[:256] for k, v in P_te.items()}
        budget = np.asarray(n_te[:256], dtype=np.int64)
        runner.prepare([workloads.config3_item(sel, i) for i in range(len(budget))])
        runner.run_prepared(budget.tolist(), seed + 99)
        runner.run_prepared(budget.tolist(), seed + 98)
        mc_rate = float(budget.sum()) / (runner.last_ms * 1e-3)
    sur = Surrogate(runner.ctx, seed=seed)
    F_tr = features(P_tr)
    keep = np.all(S_tr > 0, axis=1)           # log targets need a count in every window
    t1 = time.time()
    losses = sur.fit(F_tr[keep], S_tr[keep], epochs=epochs, batch=256, lr=1e-3, seed=seed,
                     n_hist=n_tr[keep] if weighted else None)
    train_s = time.time() - t1
    ms = sur.ens.time_infer(65536, 20)
    q_s = ms * 1e-3 / 65536
    summary = dict(
        n_train=int(keep.sum()), train_histories=int(n_tr.sum()), datasets=[str(p) for p in datasets],
        spot_datasets=[str(p) for p in spot_datasets], weighted=bool(weighted), mc_histories_per_s=mc_rate,
        train_seconds=train_s, epochs=epochs, final_member_losses=losses[-1],
        surrogate_query_s=q_s, surrogate_queries_per_s=1.0 / q_s)
    thresholds = {}
    if len(S_te):
        S_pred, spread = sur.predict(features(P_te))
        summary['held_out'] = set_metrics(S_pred, S_te, n_te)
        summary['held_out']['test_histories'] = int(n_te.sum())
        # fallback thresholds: quantiles of the members' spread of log(S90/S0)
        # on the held-out set, for fallback fractions of 0.1 %, 1 % and 5 %
        for f in FALLBACK_FRACTIONS:
            thresholds[str(f)] = float(np.quantile(spread[:, 0], 1.0 - f))
        summary['fallback_thresholds'] = thresholds
        summary['held_out']['spread_calibration'] = spread_calibration(S_pred, spread, S_te, n_te)
    if spot_datasets:
        parts = [load_spot(pth) for pth in spot_datasets]
        P_sp = _cat([a for a, _, _ in parts])
        S_sp = np.concatenate([b for _, b, _ in parts])
        n_sp = np.concatenate([c for _, _, c in parts])
        S_pred, spread = sur.predict(features(P_sp))
        summary['spot'] = spot = set_metrics(S_pred, S_sp, n_sp)
        spot['histories'] = int(n_sp.sum())
        spot.update(matched_speedup(S_pred, spot['surrogate_rms_est'], q_s, mc_rate))
        (out / 'summary.json').write_text(json.dumps(summary, indent=1))
        sur.save(out / 'weights.npz')
        log(json.dumps(summary))
        if thresholds:
            spot['fallback'] = fallback_eval(P_sp, S_pred, spread, S_sp, n_sp, thresholds, q_s, mc_rate, runner,
                                             seed + 4242)
    (out / 'summary.json').write_text(json.dumps(summary, indent=1))
    sur.save(out / 'weights.npz')
    log(json.dumps(summary))
    return summary


FALLBACK_FRACTIONS = (0.01, 0.05, 0.1, 0.2, 0.3)


def full_scale_queries(weights, n_queries=100_000, n_spot=512, counts=20_000.0, seed=49009,
                       pilot_photons=1_000_000, hi=1_000_000_000, threshold=None, log=print):
    """The surrogate leg of the full-scale workload: `n_queries` held-out
    configurations drawn afresh from the config-3 space (a seed no training or
    spot-check set used) answered by the trained ensemble in one call, and
    the first `n_spot` of them spot-checked by batched Monte Carlo with pilot
    ratio budgets for `counts` effective counts on S90/S0.  Returns the
    query times (host to host, and the inference kernel alone on the device),
    the spread distribution, the share above the fallback `threshold`, and
    the spot-check errors with the labels' own variance subtracted."""
    if not (1 <= n_spot <= n_queries):
        raise ValueError('need 1 <= n_spot <= n_queries')
    from .batched import BatchedRunner
    runner = BatchedRunner(0)
    sur = Surrogate.load(runner.ctx, weights)
    p = workloads.config3_params(n_queries, seed)
    P = {k: np.asarray(p[k], dtype=np.float64) for k in OPTICAL + TRANSPORT}
    F = features(P)
    sur.predict(F[:1024])                        # warm-up
    t0 = time.time()
    pred, spread = sur.predict(F)
    wall = time.time() - t0
    chunk = min(n_queries, 65536)               # one device-resident chunk per timed launch
    kernel_ms = sur.ens.time_infer(chunk, 20) * n_queries / chunk
    log(f'full scale: {n_queries} queries in {wall * 1e3:.1f} ms host to host, {kernel_ms:.3f} ms on the device')
    sub = {k: v[:n_spot] for k, v in P.items()}
    t1 = time.time()
    Sp = monte_carlo(sub, pilot_photons, seed + 11, runner=runner)
    budget = adaptive_budget(Sp[:, 0], pilot_photons, counts, hi=hi, s0_pilot=Sp[:, 1])
    S_mc = monte_carlo(sub, budget, seed + 13, runner=runner,
                       progress=lambda d, n: log(f'full scale: Monte Carlo spot check {d}/{n}'))
    mc_wall = time.time() - t1
    m = set_metrics(pred[:n_spot], S_mc, budget)
    out = dict(n_queries=int(n_queries), seed=int(seed), query_wall_s=wall, queries_per_s_host=n_queries / wall,
               query_kernel_ms=kernel_ms, queries_per_s_kernel=n_queries / (kernel_ms * 1e-3),
               spread_ntu_quantiles={str(q): float(np.quantile(spread[:, 0], q)) for q in (0.5, 0.9, 0.99, 0.999)},
               predicted_positive=bool(np.all(pred > 0)), n_spot=int(n_spot), spot_counts=float(counts),
               spot_histories=int(budget.sum()) + int(pilot_photons) * int(n_spot), spot_mc_wall_s=mc_wall,
               spot=m)
    if threshold is not None:
        out['fallback_threshold'] = float(threshold)
        out['fallback_fraction'] = float(np.mean(spread[:, 0] > threshold))
    return out


def evaluate_weights(weights, spot_paths, thresholds, seed=1, log=print):
    """Score trained weights on each spot-check set separately: NTU error
    against the Monte Carlo labels (label noise subtracted), the Monte Carlo
    time per query for 1 % error against the network's, and the fallback at
    each spread threshold (Monte Carlo run now).  The batched Monte Carlo
    rate is timed on the first 256 configurations of the first set."""
    from .batched import BatchedRunner
    runner = BatchedRunner(0)
    sur = Surrogate.load(runner.ctx, weights)
    q_s = sur.ens.time_infer(65536, 20) * 1e-3 / 65536
    out = dict(weights=str(weights), surrogate_query_s=q_s, surrogate_queries_per_s=1.0 / q_s, sets={})
    mc_rate = None
    for k, path in enumerate(spot_paths):
        P_sp, S_sp, n_sp = load_spot(path)
        if mc_rate is None:
            sel = {q: np.asarray(v)[:256] for q, v in P_sp.items()}
            n_mc = np.full(len(sel['rouse']), 2_000_000, dtype=np.int64)
            monte_carlo(sel, n_mc, seed + 900, runner=runner)
            monte_carlo(sel, n_mc, seed + 901, runner=runner)
            mc_rate = float(n_mc.sum()) / (runner.last_ms * 1e-3)
            out['mc_histories_per_s'] = mc_rate
        S_pred, spread = sur.predict(features(P_sp))
        m = set_metrics(S_pred, S_sp, n_sp)
        m['histories'] = int(np.sum(n_sp))
        m['distinct_optical'] = int(len({tuple(np.asarray(P_sp[q])[i] for q in OPTICAL) for i in range(len(S_sp))}))
        m.update(matched_speedup(S_pred, m['surrogate_rms_est'], q_s, mc_rate))
        if thresholds:
            m['fallback'] = fallback_eval(P_sp, S_pred, spread, S_sp, n_sp, thresholds, q_s, mc_rate, runner,
                                          seed + 4242 + k)
        out['sets'][str(path)] = m
        log(json.dumps({str(path): {q: m[q] for q in ('n', 'ntu_error_median', 'ntu_error_p90', 'label_rms',
                                                        'surrogate_rms_est', 'speedup_at_1pct')}}))
    return out


def set_metrics(S_pred, S_ref, n_ref):
    """NTU error of predictions against Monte Carlo labels that carry their
    own counting noise.  d = log of the predicted over the labelled S90/S0;
    E[d^2] = (network error)^2 + (label variance), so the network's own rms
    error is estimated as sqrt(mean d^2 - mean label variance)."""
    S_pred, S_ref = np.asarray(S_pred, dtype=np.float64), np.asarray(S_ref, dtype=np.float64)
    valid = np.all(S_ref > 0, axis=1) & np.all(S_pred > 0, axis=1)
    a, b = S_pred[valid], S_ref[valid]
    err = ntu_relative_error(a, b)
    d = np.log((a[:, 0] / a[:, 1]) / (b[:, 0] / b[:, 1]))
    v = label_variance(b, np.asarray(n_ref)[valid])[:, 0]
    return dict(n=int(valid.sum()), n_without_counts=int((~valid).sum()),
                ntu_error_median=float(np.median(err)), ntu_error_p90=float(np.percentile(err, 90)),
                ntu_error_mean=float(np.mean(err)), frac_within_1pct=float(np.mean(err <= 0.01)),
                rms_vs_labels=float(np.sqrt(np.mean(d * d))), label_rms=float(np.sqrt(np.mean(v))),
                label_se_median=float(np.median(np.sqrt(v))),
                surrogate_rms_est=float(np.sqrt(max(np.mean(d * d) - np.mean(v), 0.0))))


def spread_calibration(S_pred, spread, S_ref, n_ref, n_bins=10):
    """Estimated network rms error of log(S90/S0) per decile of the members'
    spread (label variance subtracted bin by bin)."""
    valid = np.all(S_ref > 0, axis=1)
    sp = spread[valid, 0]
    a, b = np.asarray(S_pred, dtype=np.float64)[valid], np.asarray(S_ref, dtype=np.float64)[valid]
    d2 = np.log((a[:, 0] / a[:, 1]) / (b[:, 0] / b[:, 1])) ** 2
    v = label_variance(b, np.asarray(n_ref)[valid])[:, 0]
    edges = np.quantile(sp, np.linspace(0.0, 1.0, n_bins + 1))
    rows = []
    for k in range(n_bins):
        m = (sp >= edges[k]) & ((sp < edges[k + 1]) if k < n_bins - 1 else (sp <= edges[k + 1]))
        if m.any():
            rows.append(dict(spread_lo=float(edges[k]), spread_hi=float(edges[k + 1]), n=int(m.sum()),
                             rms_est=float(np.sqrt(max(d2[m].mean() - v[m].mean(), 0.0)))))
    return rows


def matched_speedup(S_pred, rms, q_s, mc_rate):
    """Batched Monte Carlo time per query for 1 % relative standard error of
    S90/S0 (histories sized from the predicted signals), and for the
    network's own rms error, against the network's time per query."""
    n1 = mc_matched_photons(S_pred, 0.01)
    out = dict(mc_matched_photons_median=float(np.median(n1)), mc_query_s_median=float(np.median(n1) / mc_rate),
               speedup_at_1pct=float(np.median(n1) / mc_rate / q_s), meets_1pct=bool(rms <= 0.01))
    if rms > 0:
        nr = mc_matched_photons(S_pred, rms)
        out['speedup_at_matched_rms'] = float(np.median(nr) / mc_rate / q_s)
    return out


def fallback_eval(P_sp, S_pred, spread, S_ref, n_ref, thresholds, q_s, mc_rate, runner, seed, rel=0.01):
    """Answer the spot-check queries whose spread exceeds each threshold by
    batched Monte Carlo run now, histories sized for `rel` relative error from
    the network's own prediction, and score every answer against the
    spot-check labels."""
    S_pred = np.asarray(S_pred, dtype=np.float64)
    lo = min(thresholds.values())
    flag = spread[:, 0] > lo
    answers = S_pred.copy()
    n_mc = np.clip(mc_matched_photons(S_pred, rel), 200_000, 2_000_000_000).astype(np.int64)
    if flag.any():
        sub = {k: np.asarray(v)[flag] for k, v in P_sp.items()}
        answers[flag] = monte_carlo(sub, n_mc[flag], seed, runner=runner)
    t_mc = n_mc / mc_rate
    out = {}
    for name, thr in thresholds.items():
        use = spread[:, 0] > thr
        S = np.where(use[:, None], answers, S_pred)
        m = set_metrics(S, S_ref, n_ref)
        t = q_s + np.where(use, t_mc, 0.0)
        out[name] = dict(threshold=thr, fallback_fraction=float(use.mean()), ntu_error_median=m['ntu_error_median'],
                         ntu_error_p90=m['ntu_error_p90'], rms_est=m['surrogate_rms_est'],
                         frac_within_1pct=m['frac_within_1pct'],
                         speedup_vs_mc_at_1pct=float(np.mean(t_mc) / np.mean(t)))
    return out


if __name__ == '__main__':
    import argparse
    ap = argparse.ArgumentParser(description='surrogate pipeline: batched MC data, GPU training, held-out evaluation')
    ap.add_argument('--out', default='results/surrogate')
    ap.add_argument('--n-optical', type=int, default=64)
    ap.add_argument('--n-variants', type=int, default=64)
    ap.add_argument('--n-test-optical', type=int, default=16)
    ap.add_argument('--n-test-variants', type=int, default=32)
    ap.add_argument('--counts', type=float, default=200.0)
    ap.add_argument('--test-counts', type=float, default=2000.0)
    ap.add_argument('--epochs', type=int, default=60)
    ap.add_argument('--seed', type=int, default=20261204)
    ap.add_argument('--dataset', action='append', default=[], help='merge an earlier dataset.npz (repeatable)')
    ap.add_argument('--unweighted', action='store_true', help='plain MSE instead of inverse-variance weights')
    ap.add_argument('--generate-only', action='store_true', help='write dataset.npz and stop (no training)')
    ap.add_argument('--max-photons', type=float, default=2e8, help='photon budget cap per configuration')
    ap.add_argument('--adaptive-counts', type=float, default=0.0,
                    help='target 90 deg counts per training configuration from a pilot run (0: budget from --counts)')
    ap.add_argument('--ratio-budget', action='store_true',
                    help='adaptive budget for the relative error of S90/S0 rather than the 90 deg counts')
    ap.add_argument('--pilot-photons', type=int, default=1_000_000, help='pilot histories per configuration')
    ap.add_argument('--time-budget', type=float, default=None,
                    help='seconds after which training-data generation stops (configurations done so far are kept)')
    ap.add_argument('--spot-dataset', action='append', default=[],
                    help='dataset.npz whose held-out part is scored as a spot check (repeatable)')
    ap.add_argument('--adaptive-test', action='store_true',
                    help='held-out configurations get the pilot ratio budget with --test-counts')
    ap.add_argument('--evaluate', metavar='WEIGHTS',
                    help='score WEIGHTS on each --spot-dataset separately (thresholds from --thresholds) and stop')
    ap.add_argument('--full-scale', metavar='WEIGHTS',
                    help='query WEIGHTS on 1e5 fresh configurations, spot-check 512 by Monte Carlo, and stop')
    ap.add_argument('--thresholds', metavar='SUMMARY', help='summary.json whose fallback_thresholds --evaluate uses')
    ap.add_argument('--write-heldout', nargs=2, metavar=('DATASET', 'OUT'),
                    help='write the held-out part of DATASET as the benchmark held-out file OUT and stop')
    a = ap.parse_args()
    if a.write_heldout:
        print(write_heldout(*a.write_heldout))
        raise SystemExit(0)
    if a.full_scale:
        ev = Path(__file__).resolve().parent / 'surrogate_eval.json'
        thr = json.loads(ev.read_text())['bench_fallback_threshold'] if ev.exists() else None
        res = full_scale_queries(a.full_scale, seed=a.seed, threshold=thr, log=lambda s: print(s, flush=True))
        Path(a.out).mkdir(parents=True, exist_ok=True)
        (Path(a.out) / 'full_scale.json').write_text(json.dumps(res, indent=1))
        print(json.dumps({k: v for k, v in res.items() if k != 'spot'}), flush=True)
        print(json.dumps(res['spot']), flush=True)
        raise SystemExit(0)
    if a.evaluate:
        thr = json.loads(Path(a.thresholds).read_text())['fallback_thresholds'] if a.thresholds else {}
        res = evaluate_weights(a.evaluate, a.spot_dataset, thr, a.seed, log=lambda s: print(s, flush=True))
        Path(a.out).mkdir(parents=True, exist_ok=True)
        (Path(a.out) / 'evaluate.json').write_text(json.dumps(res, indent=1))
        raise SystemExit(0)
    run_pipeline(a.out, a.n_optical, a.n_variants, a.n_test_optical, a.n_test_variants, a.counts, a.test_counts,
                 a.epochs, a.seed, log=lambda s: print(s, flush=True), datasets=a.dataset, weighted=not a.unweighted,
                 generate_only=a.generate_only, hi=int(a.max_photons), spot_datasets=a.spot_dataset,
                 adaptive_counts=a.adaptive_counts, pilot_photons=a.pilot_photons, time_budget=a.time_budget,
                 ratio_budget=a.ratio_budget, adaptive_test=a.adaptive_test)
