"""Pins the FP64 surrogate-network oracle (forward and backward of the
7-64-64-64-6 hard-swish network) against the mathematics: an independent numpy
forward, and central finite differences of the loss for every parameter
block. CPU only."""
import numpy as np

from oracle import binding as o

SHAPES = [(64, 7), (64,), (64, 64), (64,), (64, 64), (64,), (6, 64), (6,)]


def unpack(p):
    out, k = [], 0
    for s in SHAPES:
        n = int(np.prod(s))
        out.append(p[k:k + n].reshape(s))
        k += n
    return out


def numpy_forward(p, x):
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    def act(z):  # hard-swish, written as x * relu6(x + 3) / 6
        return z * np.minimum(np.maximum(z + 3, 0), 6) / 6
    h = act(x @ W1.T + b1)
    h = act(h @ W2.T + b2)
    h = act(h @ W3.T + b3)
    return h @ W4.T + b4


def test_param_count_is_9222():
    assert o.mlp_param_count() == 9222 == sum(int(np.prod(s)) for s in SHAPES)


def test_forward_matches_numpy():
    rng = np.random.default_rng(0)
    p = rng.normal(0, .3, 9222)
    x = rng.normal(0, 1, (50, 7))
    assert np.allclose(o.mlp_forward(p, x), numpy_forward(p, x), rtol=1e-12, atol=1e-12)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = rng.normal(0, .3, 9222)
    x = rng.normal(0, 1, (16, 7))
    t = rng.normal(0, 1, (16, 6))
    loss, g = o.mlp_loss_grad(p, x, t)
    y = numpy_forward(p, x)
    assert abs(loss - ((y - t) ** 2).sum(1).mean()) < 1e-12
    # check a spread of indices covering every parameter block
    idx = np.concatenate([rng.choice(np.arange(a, a + int(np.prod(s))), 6, replace=False)
                          for a, s in zip(np.cumsum([0] + [int(np.prod(s)) for s in SHAPES[:-1]]), SHAPES)])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (o.mlp_loss_grad(pp, x, t)[0] - o.mlp_loss_grad(pm, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
