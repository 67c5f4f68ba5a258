"""CPU restatement of the surrogate ensemble (numpy) -- TEST INFRASTRUCTURE ONLY.

This is the checker for csrc/mlp.hip: only tests/ may import it, never the
product path.  It has no counterpart in the reference (the reference has no
surrogate); it restates the network the HIP library trains:

  n_members independent MLPs, in_dim -> hidden x n_hidden_layers -> out_dim,
  tanh-form GELU after every hidden layer, MSE loss averaged over rows and
  outputs (optionally with per-element weights), parameters flat member-major with each layer's W [out][in] then b.

With `bf16=True` it rounds to bfloat16 (round to nearest even) exactly where the
device does: every weight matrix, the activation written by each hidden layer,
and the pre-activation gradient dZ of each hidden layer.  Sums are taken in
float64, so the device (FP32 accumulation on MFMA) agrees to FP32 rounding plus
the rare BF16 tie flips that a different last bit causes.
"""
from __future__ import annotations

import numpy as np

SQ2PI = 0.7978845608028654


def bf16_round(a):
    """Round float32 values to the nearest bfloat16 (ties to even), as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(SQ2PI * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    t = np.tanh(SQ2PI * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * SQ2PI * (1.0 + 3.0 * 0.044715 * z * z)


class Shape:
    def __init__(self, n_members, in_dim, hidden, out_dim, n_hidden_layers):
        self.n_members, self.in_dim, self.hidden = n_members, in_dim, hidden
        self.out_dim, self.n_hidden_layers = out_dim, n_hidden_layers
        self.dims = [in_dim] + [hidden] * n_hidden_layers + [out_dim]

    @property
    def per_member(self):
        return sum(o * i + o for i, o in zip(self.dims[:-1], self.dims[1:]))

    def unpack(self, flat):
        """flat [n_members * per_member] -> per member list of (W, b)."""
        flat = np.asarray(flat, dtype=np.float64)
        out, off = [], 0
        for _ in range(self.n_members):
            layers = []
            for i, o in zip(self.dims[:-1], self.dims[1:]):
                W = flat[off:off + o * i].reshape(o, i)
                off += o * i
                b = flat[off:off + o]
                off += o
                layers.append((W, b))
            out.append(layers)
        return out

    def pack(self, members):
        return np.concatenate([np.concatenate([np.concatenate([W.ravel(), b]) for W, b in layers])
                               for layers in members])


def _rw(W, bf16):
    return bf16_round(W).astype(np.float64) if bf16 else W


def member_forward(layers, x, bf16=False):
    """Returns output and the cache of (input, z) per hidden layer."""
    h = np.asarray(x, dtype=np.float64)
    cache = []
    for W, b in layers[:-1]:
        z = h @ _rw(W, bf16).T + b
        y = gelu(z)
        if bf16:
            y = bf16_round(y).astype(np.float64)
        cache.append((h, z))
        h = y
    W, b = layers[-1]
    return h @ _rw(W, bf16).T + b, cache, h


def member_loss_grad(layers, x, t, bf16=False, w=None):
    """Loss sum w (out - t)^2 / (n d) (w = 1 when absent) and its gradients."""
    out, cache, h_last = member_forward(layers, x, bf16)
    n, d = out.shape
    diff = out - t
    wt = np.ones_like(diff) if w is None else np.asarray(w, dtype=np.float64)
    loss = float(np.sum(wt * diff * diff) / (n * d))
    dout = 2.0 * wt * diff / (n * d)
    grads = [None] * len(layers)
    W, b = layers[-1]
    grads[-1] = (dout.T @ h_last, dout.sum(0))
    dy = dout @ _rw(W, bf16)
    for k in range(len(layers) - 2, -1, -1):
        h_in, z = cache[k]
        dz = dy * gelu_grad(z)
        if bf16:
            dz = bf16_round(dz).astype(np.float64)
        grads[k] = (dz.T @ h_in, dz.sum(0))
        if k > 0:
            dy = dz @ _rw(layers[k][0], bf16)
    return loss, grads


def ensemble_loss_grad(shape: Shape, flat, x, t, bf16=False, w=None):
    """Per-member losses [n_members] and flat gradient (same layout as flat)."""
    losses, glist = [], []
    for layers in shape.unpack(flat):
        loss, g = member_loss_grad(layers, x, t, bf16, w)
        losses.append(loss)
        glist.append(g)
    return np.array(losses), shape.pack(glist)


def ensemble_predict(shape: Shape, flat, x, bf16=False):
    """Mean and sample standard deviation over members."""
    outs = np.stack([member_forward(layers, x, bf16)[0] for layers in shape.unpack(flat)])
    mean = outs.mean(0)
    std = outs.std(0, ddof=1) if shape.n_members > 1 else np.zeros_like(mean)
    return mean, std
