"""The CPU oracle of the warm-start network (oracle/nn_oracle.py) pinned by
mathematics: its backward pass equals central differences of its forward
loss in FP64 (a transposed weight, a wrong tap offset or a missing SiLU'
fails this), the transposed convolution is the adjoint of the convolution,
and its BF16 rounding agrees with torch's float -> bfloat16 conversion
(round to nearest, ties to even)."""

import numpy as np
import pytest

from oracle import nn_oracle as nn


def _net(rng, n_in=5, layers=3, width=32):
    L = []
    for k in range(layers):
        fan = 27 * (n_in if k == 0 else width)
        w = rng.normal(0, np.sqrt(2.0 / fan), size=(27, width, width))
        if k == 0:
            w[:, :, n_in:] = 0
        L.append((w, rng.normal(0, 0.1, size=width)))
    return L


def test_backward_matches_central_differences():
    rng = np.random.default_rng(0)
    x = np.zeros((2, 3, 4, 5, 32))
    x[..., :5] = rng.normal(size=(2, 3, 4, 5, 5))
    layers = _net(rng)
    target = np.zeros((2, 3, 4, 5, 16))
    target[..., :4] = rng.normal(size=(2, 3, 4, 5, 4))
    target[..., 15] = (rng.uniform(size=(2, 3, 4, 5)) > 0.3)
    chw = np.array([1.0, 2.0, 0.5, 1.5] + [0.0] * 12)
    inv = 1.0 / 37.0

    def total(ls):
        y, _, _ = nn.network_forward(x, ls)
        return nn.loss(y, target, chw, 4, inv)[0]

    y, acts, zs = nn.network_forward(x, layers)
    _, dy = nn.loss(y, target, chw, 4, inv)
    grads = nn.network_backward(dy, layers, acts, zs)
    h = 1e-6
    for k in range(3):
        for (t, o, i) in ((0, 0, 0), (13, 5, 3), (26, 31, 4), (7, 2, 20) if k else (7, 2, 1)):
            ls = [(w.copy(), b.copy()) for w, b in layers]
            ls[k][0][t, o, i] += h
            fp = total(ls)
            ls[k][0][t, o, i] -= 2 * h
            fm = total(ls)
            fd = (fp - fm) / (2 * h)
            assert grads[k][0][t, o, i] == pytest.approx(fd, rel=1e-5, abs=1e-9), (k, t, o, i)
        for o in (0, 17, 31):
            ls = [(w.copy(), b.copy()) for w, b in layers]
            ls[k][1][o] += h
            fp = total(ls)
            ls[k][1][o] -= 2 * h
            fm = total(ls)
            assert grads[k][1][o] == pytest.approx((fp - fm) / (2 * h), rel=1e-5, abs=1e-9)


def test_transposed_convolution_is_the_adjoint():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(1, 4, 3, 5, 32))
    dz = rng.normal(size=(1, 4, 3, 5, 32))
    w = rng.normal(size=(27, 32, 32))
    lhs = np.sum(dz * nn.conv(x, w))
    rhs = np.sum(x * nn.conv_backward_data(dz, w))
    assert lhs == pytest.approx(rhs, rel=1e-12)
    # a single tap moves values by its offset
    w1 = np.zeros((27, 32, 32))
    w1[nn.TAPS.index((1, 0, -1)), 3, 7] = 1.0
    out = nn.conv(x, w1)
    np.testing.assert_array_equal(out[0, 1, :, 2, 3], x[0, 2, :, 1, 7])
    assert np.all(out[0, 3, :, :, 3] == 0) and np.all(out[0, :, :, 0, 3] == 0)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(2)
    v = np.concatenate([rng.normal(size=2000) * 10.0 ** rng.integers(-6, 6, 2000),
                        [0.0, -0.0, 1.0, 1.00390625, 1.01171875, 3.0e38, -2.5e-40]])
    want = torch.from_numpy(v.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(nn.bf16(v), want)
