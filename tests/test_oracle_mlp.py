"""Pins of the surrogate's CPU oracle (oracle/mlp_oracle.py).

The oracle's forward pass and gradients are checked against torch autograd in
float64 (an independent implementation of the same network: Linear layers,
tanh-form GELU, MSE), and its gradient against central finite differences.
The BF16 rounding helper is checked against torch's own float32 -> bfloat16
conversion.  No GPU.
"""
import numpy as np
import pytest

from oracle import mlp_oracle as mo

torch = pytest.importorskip('torch')


def _random_params(shape, seed):
    g = np.random.default_rng(seed)
    return g.normal(0.0, 0.3, shape.n_members * shape.per_member)


def _torch_member(layers, x, t):
    ps, h = [], torch.tensor(x, dtype=torch.float64)
    for k, (W, b) in enumerate(layers):
        Wt = torch.tensor(W, dtype=torch.float64, requires_grad=True)
        bt = torch.tensor(b, dtype=torch.float64, requires_grad=True)
        ps += [Wt, bt]
        h = h @ Wt.T + bt
        if k < len(layers) - 1:
            h = torch.nn.functional.gelu(h, approximate='tanh')
    loss = torch.mean((h - torch.tensor(t, dtype=torch.float64)) ** 2)
    loss.backward()
    return float(loss.detach()), h.detach().numpy(), [p.grad.numpy() for p in ps]


def test_oracle_matches_torch_autograd():
    shape = mo.Shape(3, 8, 32, 3, 3)
    flat = _random_params(shape, 1)
    g = np.random.default_rng(2)
    x = g.normal(size=(17, 8))
    t = g.normal(size=(17, 3))
    losses, grad = mo.ensemble_loss_grad(shape, flat, x, t, bf16=False)
    members = shape.unpack(flat)
    glist = shape.unpack(grad)
    for m, layers in enumerate(members):
        loss_t, out_t, grads_t = _torch_member(layers, x, t)
        out_o = mo.member_forward(layers, x)[0]
        np.testing.assert_allclose(out_o, out_t, rtol=1e-12, atol=1e-12)
        assert abs(losses[m] - loss_t) <= 1e-12 * max(1.0, loss_t)
        mine = [a for W, b in glist[m] for a in (W, b)]
        for a, b in zip(mine, grads_t):
            np.testing.assert_allclose(a, b, rtol=1e-10, atol=1e-12)


def test_oracle_gradient_finite_differences():
    shape = mo.Shape(1, 3, 4, 2, 2)
    flat = _random_params(shape, 3)
    g = np.random.default_rng(4)
    x = g.normal(size=(5, 3))
    t = g.normal(size=(5, 2))
    _, grad = mo.ensemble_loss_grad(shape, flat, x, t)
    h = 1e-6
    num = np.empty_like(flat)
    for i in range(flat.size):
        e = np.zeros_like(flat)
        e[i] = h
        lp = mo.ensemble_loss_grad(shape, flat + e, x, t)[0][0]
        lm = mo.ensemble_loss_grad(shape, flat - e, x, t)[0][0]
        num[i] = (lp - lm) / (2 * h)
    np.testing.assert_allclose(grad, num, rtol=1e-6, atol=1e-8)


def test_bf16_rounding_matches_torch():
    g = np.random.default_rng(5)
    a = np.concatenate([g.normal(size=4000).astype(np.float32) * 10.0 ** g.integers(-6, 6, 4000),
                        np.array([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -0.0, 65504.0], dtype=np.float32)])
    mine = mo.bf16_round(a)
    ref = torch.tensor(a).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(mine.view(np.uint32), ref.view(np.uint32))


def test_ensemble_statistics():
    shape = mo.Shape(4, 2, 8, 1, 1)
    flat = _random_params(shape, 6)
    x = np.random.default_rng(7).normal(size=(9, 2))
    mean, std = mo.ensemble_predict(shape, flat, x)
    outs = np.stack([mo.member_forward(l, x)[0] for l in shape.unpack(flat)])
    np.testing.assert_allclose(mean, outs.mean(0), rtol=1e-14)
    np.testing.assert_allclose(std, outs.std(0, ddof=1), rtol=1e-12)


def test_weighted_oracle_matches_torch_and_finite_differences():
    shape = mo.Shape(2, 4, 16, 3, 2)
    flat = _random_params(shape, 8)
    g = np.random.default_rng(9)
    x = g.normal(size=(11, 4))
    t = g.normal(size=(11, 3))
    w = g.uniform(0.0, 3.0, size=(11, 3))
    losses, grad = mo.ensemble_loss_grad(shape, flat, x, t, w=w)
    glist = shape.unpack(grad)
    for m, layers in enumerate(shape.unpack(flat)):
        ps, h = [], torch.tensor(x, dtype=torch.float64)
        for k, (W, b) in enumerate(layers):
            Wt = torch.tensor(W, dtype=torch.float64, requires_grad=True)
            bt = torch.tensor(b, dtype=torch.float64, requires_grad=True)
            ps += [Wt, bt]
            h = h @ Wt.T + bt
            if k < len(layers) - 1:
                h = torch.nn.functional.gelu(h, approximate='tanh')
        loss = torch.mean(torch.tensor(w) * (h - torch.tensor(t)) ** 2)
        loss.backward()
        assert abs(losses[m] - float(loss)) <= 1e-12 * max(1.0, float(loss))
        mine = [a for W, b in glist[m] for a in (W, b)]
        for a, b in zip(mine, [p.grad.numpy() for p in ps]):
            np.testing.assert_allclose(a, b, rtol=1e-10, atol=1e-12)
    # unit weights reduce to the plain MSE
    l1, g1 = mo.ensemble_loss_grad(shape, flat, x, t, w=np.ones_like(w))
    l0, g0 = mo.ensemble_loss_grad(shape, flat, x, t)
    np.testing.assert_array_equal(l1, l0)
    np.testing.assert_array_equal(g1, g0)
    # finite differences on one member
    one = mo.Shape(1, 4, 16, 3, 2)
    f1 = flat[:one.per_member]
    _, gr = mo.ensemble_loss_grad(one, f1, x, t, w=w)
    h = 1e-6
    for i in np.random.default_rng(10).choice(f1.size, 40, replace=False):
        e = np.zeros_like(f1)
        e[i] = h
        num = (mo.ensemble_loss_grad(one, f1 + e, x, t, w=w)[0][0] - mo.ensemble_loss_grad(one, f1 - e, x, t, w=w)[0][0]) / (2 * h)
        assert abs(gr[i] - num) <= 1e-6 * max(1.0, abs(num)), (i, gr[i], num)
