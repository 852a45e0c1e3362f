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


def test_conv3_matches_direct_sum():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((4, 3, 5, 16))
    W = rng.standard_normal((27, 16, 16))
    y = cn.conv(x, W)
    for (i, j, k, co) in ((0, 2, 4, 5), (2, 1, 0, 11), (3, 0, 2, 0)):
        ref = 0.0
        for tap in range(27):
            ii, jj, kk = i + tap // 9 - 1, j + (tap // 3) % 3 - 1, k + tap % 3 - 1
            if 0 <= ii < 4 and 0 <= jj < 3 and 0 <= kk < 5:
                ref += x[ii, jj, kk] @ W[tap, co]
        assert abs(y[i, j, k, co] - ref) < 1e-12


def test_pool_up_adjoint_3d():
    rng = np.random.default_rng(4)
    x = rng.standard_normal((7, 5, 3, 3))
    g = rng.standard_normal((4, 3, 2, 3))
    assert abs(np.sum(cn.pool(x) * g) - np.sum(x * cn.pool_grad(g, 7, 5, 3))) < 1e-12
    f = rng.standard_normal((7, 5, 3, 3))
    assert abs(np.sum(cn.up(g, 7, 5, 3) * f) - np.sum(g * cn.up_grad(f, 4, 3, 2))) < 1e-12
    assert np.allclose(cn.pool(np.ones((7, 5, 3, 2))), 1.0)
    # a 2 x 2 x 2 block averages its eight values
    assert abs(cn.pool(x)[0, 0, 0, 1] - x[:2, :2, :2, 1].mean()) < 1e-14


def test_backward_finite_differences_3d():
    rng = np.random.default_rng(5)
    shape = (5, 4, 3)
    params = rng.standard_normal(cn.num_params(3)) * 0.15
    x0 = rng.standard_normal(shape + (16,))
    gy = rng.standard_normal(shape + (16,))
    y, A, Ws = cn.forward(x0, params)
    g = cn.flatten(cn.backward(gy, A, Ws))
    assert len(g) == cn.num_params(3)
    idx = rng.choice(len(params), 30, replace=False)
    h = 1e-6
    for q in idx:
        pp, pm = params.copy(), params.copy()
        pp[q] += h
        pm[q] -= h
        fd = (np.sum(cn.forward(x0, pp)[0] * gy) - np.sum(cn.forward(x0, pm)[0] * gy)) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 * max(1.0, abs(fd)), (q, fd, g[q])
