"""Pins of the CPU restatement of the preconditioner network
(oracle/convnet.py): its backward pass against central finite differences
of its FP64 forward, and the building blocks against direct definitions."""
import numpy as np

from oracle import convnet as cn


def test_conv_matches_direct_sum():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((5, 4, 16))
    W = rng.standard_normal((9, 16, 16))
    y = cn.conv(x, W)
    i, j, co = 0, 3, 5
    ref = 0.0
    for tap in range(9):
        ii, jj = i + tap // 3 - 1, j + tap % 3 - 1
        if 0 <= ii < 5 and 0 <= jj < 4:
            ref += x[ii, jj] @ W[tap, co]
    assert abs(y[i, j, co] - ref) < 1e-12


def test_pool_up_adjoint():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((7, 5, 3))
    g = rng.standard_normal((4, 3, 3))
    # <pool(x), g> = <x, pool_grad(g)> and <up(c), f> = <c, up_grad(f)>
    assert abs(np.sum(cn.pool(x) * g) - np.sum(x * cn.pool_grad(g, 7, 5))) < 1e-12
    f = rng.standard_normal((7, 5, 3))
    assert abs(np.sum(cn.up(g, 7, 5) * f) - np.sum(g * cn.up_grad(f, 4, 3))) < 1e-12
    assert np.allclose(cn.pool(np.ones((7, 5, 2))), 1.0)


def test_backward_finite_differences():
    rng = np.random.default_rng(2)
    H, W = 9, 7
    params = rng.standard_normal(cn.num_params()) * 0.2
    x0 = rng.standard_normal((H, W, 16))
    gy = rng.standard_normal((H, W, 16))
    y, A, Ws = cn.forward(x0, params)
    g = cn.flatten(cn.backward(gy, A, Ws))
    idx = rng.choice(len(params), 40, replace=False)
    h = 1e-6
    for q in idx:
        pp, pm = params.copy(), params.copy()
        pp[q] += h
        pm[q] -= h
        fd = (np.sum(cn.forward(x0, pp)[0] * gy) - np.sum(cn.forward(x0, pm)[0] * gy)) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 * max(1.0, abs(fd)), (q, fd, g[q])
