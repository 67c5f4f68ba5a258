"""Learned surrogate of the detector signals (config 4; no reference counterpart).

An ensemble of MLPs (default 5 members, 8 -> 256 -> 256 -> 256 -> 3, GELU) maps
the eight configuration parameters

    log10 concentration, PSD median (um), sigma_g, m_r, log10 m_i,
    wavelength (nm), vortex depth (mm), Rouse number

to the 90 deg, 0 deg and 135 deg detector signals (symmetrised hardware hit
fractions of launched histories).  It is trained on the GPU by the library's own
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

FEATURES = ('log10_conc_mg_l', 'median_um', 'sigma_g', 'm_r', 'log10_m_i', 'wavelength_nm', 'vortex_depth_mm',
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

    def grad(self, x, y):
        x, y = _fp(x), _fp(y)
        loss = np.empty(self.n_members, dtype=np.float32)
        g = np.empty(self.n_params, dtype=np.float32)
        self.ctx.check(self.ctx.lib.clr_mlp_grad(self.ctx.h, x.ctypes.data_as(P(C.c_float)),
                                                 y.ctypes.data_as(P(C.c_float)), len(x),
                                                 loss.ctypes.data_as(P(C.c_float)), g.ctypes.data_as(P(C.c_float))),
                       'clr_mlp_grad')
        return loss, g

    def train(self, x, y, batch=512, epochs=10, lr=1e-3, seed=0):
        x, y = _fp(x), _fp(y)
        loss = np.empty(self.n_members, dtype=np.float32)
        self.ctx.check(self.ctx.lib.clr_mlp_train(self.ctx.h, x.ctypes.data_as(P(C.c_float)),
                                                  y.ctypes.data_as(P(C.c_float)), len(x), int(batch), int(epochs),
                                                  C.c_float(lr), C.c_uint64(seed), loss.ctypes.data_as(P(C.c_float))),
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
    return np.stack([np.log10(p['conc_mg_l']), p['median_um'], p['sigma_g'], p['m_r'], np.log10(p['m_i']),
                     p['wavelength_nm'], p['vortex_depth_mm'], p['rouse']], axis=1).astype(np.float64)


def signals_from_tallies(T, n_histories):
    """S90, S0, S135: symmetrised hardware hits / launched histories."""
    det = T['det'].sum(0)   # [4][18]
    aux = T['aux'].sum(0)   # [2][n_aux]
    s90 = 0.5 * (det[0, 9] + det[1, 9]) / n_histories
    s0 = 0.5 * (det[0, 0] + det[1, 0]) / n_histories
    s135 = 0.5 * (aux[0, 0] + aux[1, 0]) / n_histories
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


class Surrogate:
    """Feature/target transforms around an MlpEnsemble, with MC fallback."""

    def __init__(self, ctx: Context, n_members=5, hidden=256, n_hidden_layers=3, seed=0):
        self.ens = MlpEnsemble(ctx, n_members, len(FEATURES), hidden, len(SIGNALS), n_hidden_layers, seed)
        self.norm: Normaliser | None = None

    def _x(self, F):
        return (F - self.norm.x_mean) / self.norm.x_std

    def fit(self, F, S, epochs=40, batch=512, lr=1e-3, seed=0, schedule=(1.0, 0.3, 0.1)):
        """Train on raw features F [n][8] and signals S [n][3] (log targets).
        The learning rate steps down over `schedule` (equal shares of epochs)."""
        Y = np.log(np.maximum(S, FLOOR))
        self.norm = Normaliser(F.mean(0), F.std(0) + 1e-12, Y.mean(0), Y.std(0) + 1e-12)
        x, y = self._x(F), (Y - self.norm.y_mean) / self.norm.y_std
        losses = []
        per = max(1, epochs // len(schedule))
        for k, f in enumerate(schedule):
            losses.append(self.ens.train(x, y, batch, per, lr * f, seed + k).tolist())
        return losses

    def predict(self, F):
        """Signals [n][3] and the members' spread of log S [n][3] (relative)."""
        m, s = self.ens.predict(self._x(F))
        logS = m * self.norm.y_std + self.norm.y_mean
        return np.exp(logS), s * self.norm.y_std

    def query(self, params, threshold=0.01, mc_photons=10_000_000, seed=1, runner=None):
        """Signals for a parameter dict: network where the ensemble spread of
        log(S90/S0) is within `threshold`, Monte Carlo elsewhere.  Returns
        (signals, used_mc mask)."""
        F = features(params)
        S, spread = self.predict(F)
        ratio_spread = np.sqrt(spread[:, 0] ** 2 + spread[:, 1] ** 2)
        mc = ratio_spread > threshold
        if np.any(mc):
            sub = {k: np.asarray(v)[mc] for k, v in params.items()}
            S[mc] = monte_carlo(sub, mc_photons, seed, runner=runner)
        return S, mc


def ntu_relative_error(S_pred, S_ref):
    """|dNTU|/NTU with NTU = K S90/S0."""
    return np.abs((S_pred[:, 0] / S_pred[:, 1]) / (S_ref[:, 0] / S_ref[:, 1]) - 1.0)
