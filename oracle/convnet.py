"""CPU restatement of the learned preconditioner's network -- TEST
INFRASTRUCTURE ONLY (the checker of the HIP kernels in
pansfem2_amd/csrc/pf_learned.hip; nothing in the product imports it).

The network has no counterpart in the reference.  It is restated here in
numpy, layer for layer, in the same order and with the same tensor layout
(H x W x C, channels fastest) as the device code:

  x0 -> c0a -> c0b = e0 -> pool -> c1a -> c1b = e1 -> pool -> c2a -> c2b = e2
     -> pool -> c3a -> c3b = b3 -> up + e2 -> d2 -> up + e1 -> d1 -> up + e0
     -> d0 -> out (linear)

3x3 (quad4) or 3x3x3 (hex8) convolutions with zero padding, weights
[tap][cout][cin], tap = 3 (di + 1) + (dj + 1) or 9 (di + 1) + 3 (dj + 1) +
(dk + 1); ReLU after every conv but the last; 2x2 (x2) average pooling with
ceil (partial blocks divide by their own count); nearest upsampling plus skip
addition.  `bf16=True` rounds every stored activation
to bfloat16 (round to nearest even) as the device does; the backward pass is
the exact FP64 gradient of the FP64 network (bf16=False), pinned by finite
differences in tests/test_convnet_oracle.py.
"""
import itertools

import numpy as np

SPEC = [(16, 16, 0, 1), (16, 16, 0, 1), (16, 32, 1, 1), (32, 32, 1, 1), (32, 32, 2, 1), (32, 32, 2, 1),
        (32, 64, 3, 1), (64, 32, 3, 1), (32, 32, 2, 1), (32, 16, 1, 1), (16, 16, 0, 1), (16, 16, 0, 0)]


def to_bf16(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def taps(dim):
    """Tap offsets in device order: lexicographic over (d0, d1[, d2]) in
    {-1, 0, 1}, so tap = 3 (di + 1) + (dj + 1) in 2D and
    9 (di + 1) + 3 (dj + 1) + (dk + 1) in 3D."""
    return list(itertools.product((-1, 0, 1), repeat=dim))


def split_params(flat, dim=2):
    out, off = [], 0
    nt = 3 ** dim
    for cin, cout, _, _ in SPEC:
        n = nt * cout * cin
        out.append(np.asarray(flat[off:off + n], dtype=np.float64).reshape(nt, cout, cin))
        off += n
    return out


def _shift(shape, off):
    """Slices of the one-padded array that hold x[p + off] for every p."""
    return tuple(slice(1 + o, 1 + o + n) for o, n in zip(off, shape)) + (slice(None),)


def conv(x, W):
    """y[p, co] = sum_{tap, ci} x[p + off(tap), ci] W[tap, co, ci] (zero padding);
    p runs over the 2 or 3 spatial axes (channels last)."""
    sp = x.shape[:-1]
    xp = np.pad(x, [(1, 1)] * len(sp) + [(0, 0)])
    y = np.zeros(sp + (W.shape[1],))
    for t, off in enumerate(taps(len(sp))):
        y += xp[_shift(sp, off)] @ W[t].T
    return y


def conv_grads(x, gz, W):
    """(dL/dx, dL/dW) of y = conv(x, W) for dL/dy = gz."""
    sp = x.shape[:-1]
    xp = np.pad(x, [(1, 1)] * len(sp) + [(0, 0)])
    gW = np.zeros_like(W)
    gxp = np.zeros_like(xp)
    for t, off in enumerate(taps(len(sp))):
        xs = xp[_shift(sp, off)]
        gW[t] = np.tensordot(gz, xs, axes=(list(range(len(sp))), list(range(len(sp)))))
        gxp[_shift(sp, off)] += gz @ W[t]
    inner = tuple(slice(1, -1) for _ in sp) + (slice(None),)
    return gxp[inner], gW


def _blocks(dim):
    return list(itertools.product((0, 1), repeat=dim))


def pool(x):
    """2 x 2 (x 2) average pooling with ceil; partial blocks divide by their
    own count."""
    sp, C = x.shape[:-1], x.shape[-1]
    cs = tuple((n + 1) // 2 for n in sp)
    s = np.zeros(cs + (C,))
    cnt = np.zeros(cs + (1,))
    for blk_off in _blocks(len(sp)):
        blk = x[tuple(slice(o, None, 2) for o in blk_off)]
        idx = tuple(slice(0, m) for m in blk.shape[:-1])
        s[idx] += blk
        cnt[idx] += 1
    return s / cnt


def pool_grad(g, *shape):
    cs = g.shape[:-1]
    cnt = np.zeros(cs + (1,))
    for blk_off in _blocks(len(cs)):
        cnt[tuple(slice(0, len(range(o, n, 2))) for o, n in zip(blk_off, shape))] += 1
    gf = g / cnt
    for ax in range(len(cs)):
        gf = np.repeat(gf, 2, ax)
    return gf[tuple(slice(0, n) for n in shape)]


def up(x, *shape):
    for ax in range(x.ndim - 1):
        x = np.repeat(x, 2, ax)
    return x[tuple(slice(0, n) for n in shape)]


def up_grad(g, *cshape):
    out = np.zeros(tuple(cshape) + (g.shape[-1],))
    for blk_off in _blocks(g.ndim - 1):
        blk = g[tuple(slice(o, None, 2) for o in blk_off)]
        out[tuple(slice(0, m) for m in blk.shape[:-1])] += blk
    return out


def forward(x0, params, bf16=False):
    """Returns y (spatial x 16) and the saved activations; x0 is H x W x 16
    (quad4 nodes) or H x W x D x 16 (hex8 nodes)."""
    Ws = split_params(params, x0.ndim - 1) if not isinstance(params, list) else params
    if bf16:
        Ws = [to_bf16(w) for w in Ws]
        x0 = to_bf16(x0)
    q = to_bf16 if bf16 else (lambda a: a)
    relu = lambda a: np.maximum(a, 0.0)
    A = {"x0": x0}
    A["a0"] = q(relu(conv(x0, Ws[0])))
    A["e0"] = q(relu(conv(A["a0"], Ws[1])))
    A["x1"] = q(pool(A["e0"]))
    A["a1"] = q(relu(conv(A["x1"], Ws[2])))
    A["e1"] = q(relu(conv(A["a1"], Ws[3])))
    A["x2"] = q(pool(A["e1"]))
    A["a2"] = q(relu(conv(A["x2"], Ws[4])))
    A["e2"] = q(relu(conv(A["a2"], Ws[5])))
    A["x3"] = q(pool(A["e2"]))
    A["a3"] = q(relu(conv(A["x3"], Ws[6])))
    A["b3"] = q(relu(conv(A["a3"], Ws[7])))
    A["u2"] = q(up(A["b3"], *A["e2"].shape[:-1]) + A["e2"])
    A["d2"] = q(relu(conv(A["u2"], Ws[8])))
    A["u1"] = q(up(A["d2"], *A["e1"].shape[:-1]) + A["e1"])
    A["d1"] = q(relu(conv(A["u1"], Ws[9])))
    A["u0"] = q(up(A["d1"], *A["e0"].shape[:-1]) + A["e0"])
    A["d0"] = q(relu(conv(A["u0"], Ws[10])))
    y = conv(A["d0"], Ws[11])
    return y, A, Ws


def backward(gy, A, Ws):
    """Gradients (list per layer, [tap][cout][cin]) of sum(y * gy)."""
    g = [None] * 12
    mask = lambda a: (a > 0).astype(np.float64)
    gd0, g[11] = conv_grads(A["d0"], gy, Ws[11])
    gu0, g[10] = conv_grads(A["u0"], gd0 * mask(A["d0"]), Ws[10])
    ge0 = gu0.copy()
    gd1 = up_grad(gu0, *A["d1"].shape[:-1])
    gu1, g[9] = conv_grads(A["u1"], gd1 * mask(A["d1"]), Ws[9])
    ge1 = gu1.copy()
    gd2 = up_grad(gu1, *A["d2"].shape[:-1])
    gu2, g[8] = conv_grads(A["u2"], gd2 * mask(A["d2"]), Ws[8])
    ge2 = gu2.copy()
    gb3 = up_grad(gu2, *A["b3"].shape[:-1])
    ga3, g[7] = conv_grads(A["a3"], gb3 * mask(A["b3"]), Ws[7])
    gx3, g[6] = conv_grads(A["x3"], ga3 * mask(A["a3"]), Ws[6])
    ge2 = ge2 + pool_grad(gx3, *A["e2"].shape[:-1])
    ga2, g[5] = conv_grads(A["a2"], ge2 * mask(A["e2"]), Ws[5])
    gx2, g[4] = conv_grads(A["x2"], ga2 * mask(A["a2"]), Ws[4])
    ge1 = ge1 + pool_grad(gx2, *A["e1"].shape[:-1])
    ga1, g[3] = conv_grads(A["a1"], ge1 * mask(A["e1"]), Ws[3])
    gx1, g[2] = conv_grads(A["x1"], ga1 * mask(A["a1"]), Ws[2])
    ge0 = ge0 + pool_grad(gx1, *A["e0"].shape[:-1])
    ga0, g[1] = conv_grads(A["a0"], ge0 * mask(A["e0"]), Ws[1])
    _, g[0] = conv_grads(A["x0"], ga0 * mask(A["a0"]), Ws[0])
    return g


def flatten(gs):
    return np.concatenate([x.reshape(-1) for x in gs])


def num_params(dim=2):
    return sum(3 ** dim * cout * cin for cin, cout, _, _ in SPEC)
