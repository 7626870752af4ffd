"""The warm-start network's HIP kernels (include/vlbm_nn.h) against the CPU
FP64 oracle (oracle/nn_oracle.py) applied to the kernels' own BF16 inputs:
forward convolution (hidden and output epilogues), backward-data, weight and
bias gradients, loss, weight packing and Adam, each on an odd-sized grid so
that partial voxel blocks are exercised; then the whole network's
gradients, and training that lowers the loss.

Bars: FP32 results within 2e-6 of the FP64 value relative to the largest
magnitude (FP32 accumulation of 864 products); BF16 results equal to the
oracle rounded to BF16, allowing one BF16 step where the FP32 and FP64 sums
straddle a rounding boundary (at most 1% of the values)."""

import ctypes as C

import numpy as np
import pytest

from oracle import nn_oracle as nn

pytestmark = pytest.mark.gpu

SHAPE = (2, 5, 6, 7)  # batch, nx, ny, nz


def _setup():
    torch = pytest.importorskip("torch")
    from curved_lbm import _native

    return torch, _native, _native.load_library()


def _padded(torch, x, dtype):
    """(B, X, Y, Z, C) -> zero-halo device tensor (B, X+2, Y+2, Z+2, C)."""
    B, X, Y, Z, Cc = x.shape
    out = np.zeros((B, X + 2, Y + 2, Z + 2, Cc))
    out[:, 1:-1, 1:-1, 1:-1] = x
    return torch.from_numpy(out).to(dtype=dtype, device="cuda")


def _interior(t):
    return t.double().cpu().numpy()[:, 1:-1, 1:-1, 1:-1]


def _bf16_close(got, want_exact):
    """got (BF16 values) vs the BF16 rounding of the FP64 value."""
    want = nn.bf16(want_exact)
    same = got == want
    assert same.mean() > 0.99, same.mean()
    ulp = np.abs(want) * 2.0 ** -7 + 1e-30
    assert np.all(np.abs(got - want) <= ulp * 1.01)


def _rand_bf16(rng, shape, scale=1.0):
    return nn.bf16(rng.normal(0, scale, size=shape))


def test_gpu_conv_forward_matches_oracle():
    torch, _native, lib = _setup()
    rng = np.random.default_rng(0)
    x = _rand_bf16(rng, SHAPE + (32,))
    w = _rand_bf16(rng, (27, 32, 32), 0.05)
    b = rng.normal(0, 0.1, 32).astype(np.float32).astype(np.float64)
    g = _native.NNGrid(*SHAPE)
    xd = _padded(torch, x, torch.bfloat16)
    wd = torch.from_numpy(w.reshape(-1)).to(dtype=torch.bfloat16, device="cuda")
    bd = torch.from_numpy(b).to(dtype=torch.float32, device="cuda")
    z = torch.zeros_like(xd)
    a = torch.zeros_like(xd)
    y = torch.zeros(xd.shape, dtype=torch.float32, device="cuda")
    assert lib.vlbm_nn_conv_forward(C.byref(g), xd.data_ptr(), wd.data_ptr(), bd.data_ptr(), 0, z.data_ptr(),
                                    a.data_ptr(), None, None) == 0
    assert lib.vlbm_nn_conv_forward(C.byref(g), xd.data_ptr(), wd.data_ptr(), bd.data_ptr(), 1, None, None,
                                    y.data_ptr(), None) == 0
    torch.cuda.synchronize()
    s = nn.conv(x, w, b)
    got_y = _interior(y)
    assert np.max(np.abs(got_y - s)) <= 2e-6 * np.max(np.abs(s))
    _bf16_close(_interior(z), s)
    _bf16_close(_interior(a), nn.silu(s))
    # the halo stays zero
    assert float(z.float().abs()[:, 0].max()) == 0.0 and float(y.abs()[:, :, :, -1].max()) == 0.0


def test_gpu_conv_backward_data_matches_oracle():
    torch, _native, lib = _setup()
    rng = np.random.default_rng(1)
    dz = _rand_bf16(rng, SHAPE + (32,))
    zp = _rand_bf16(rng, SHAPE + (32,), 2.0)
    master = rng.normal(0, 0.05, (27, 32, 32)).astype(np.float32)
    g = _native.NNGrid(*SHAPE)
    md = torch.from_numpy(master.reshape(-1)).cuda()
    w = torch.zeros(27 * 32 * 32, dtype=torch.bfloat16, device="cuda")
    wt = torch.zeros_like(w)
    assert lib.vlbm_nn_pack_weights(md.data_ptr(), w.data_ptr(), wt.data_ptr(), None) == 0
    dzd = _padded(torch, dz, torch.bfloat16)
    zpd = _padded(torch, zp, torch.bfloat16)
    out = torch.zeros_like(dzd)
    assert lib.vlbm_nn_conv_backward_data(C.byref(g), dzd.data_ptr(), wt.data_ptr(), zpd.data_ptr(),
                                          out.data_ptr(), None) == 0
    torch.cuda.synchronize()
    wb = nn.bf16(master)
    np.testing.assert_array_equal(w.float().cpu().numpy().reshape(27, 32, 32), wb)
    np.testing.assert_array_equal(wt.float().cpu().numpy().reshape(27, 32, 32),
                                  wb[::-1].transpose(0, 2, 1))
    _bf16_close(_interior(out), nn.conv_backward_data(dz, wb) * nn.dsilu(zp))


def test_gpu_conv_weight_gradient_matches_oracle():
    torch, _native, lib = _setup()
    rng = np.random.default_rng(2)
    x = _rand_bf16(rng, SHAPE + (32,))
    dz = _rand_bf16(rng, SHAPE + (32,))
    g = _native.NNGrid(*SHAPE)
    xd, dzd = _padded(torch, x, torch.bfloat16), _padded(torch, dz, torch.bfloat16)
    for nparts in (1, 7, 1024):
        part = torch.zeros((nparts, 27 * 1024 + 32), dtype=torch.float32, device="cuda")
        dw = torch.zeros(27 * 1024, dtype=torch.float32, device="cuda")
        db = torch.zeros(32, dtype=torch.float32, device="cuda")
        assert lib.vlbm_nn_conv_backward_weight(C.byref(g), xd.data_ptr(), dzd.data_ptr(), part.data_ptr(), nparts,
                                                dw.data_ptr(), db.data_ptr(), None) == 0
        torch.cuda.synchronize()
        want_w, want_b = nn.conv_backward_weight(x, dz)
        got_w = dw.cpu().numpy().astype(np.float64).reshape(27, 32, 32)
        assert np.max(np.abs(got_w - want_w)) <= 2e-6 * np.max(np.abs(want_w)), nparts
        assert np.max(np.abs(db.cpu().numpy() - want_b)) <= 2e-6 * np.max(np.abs(want_b)), nparts


def test_gpu_loss_matches_oracle():
    torch, _native, lib = _setup()
    rng = np.random.default_rng(3)
    y = rng.normal(size=SHAPE + (32,)).astype(np.float32).astype(np.float64)
    t = np.zeros(SHAPE + (16,))
    t[..., :12] = rng.normal(size=SHAPE + (12,)).astype(np.float32)
    t[..., 15] = rng.uniform(size=SHAPE) > 0.4
    chw = np.zeros(32)
    chw[:12] = rng.uniform(0.5, 2.0, 12).astype(np.float32)
    inv = np.float32(1.0 / 123.0)
    g = _native.NNGrid(*SHAPE)
    yd = _padded(torch, y, torch.float32)
    td = _padded(torch, t, torch.float32)
    cd = torch.from_numpy(chw).to(dtype=torch.float32, device="cuda")
    dz = torch.full(yd.shape, 7.0, dtype=torch.bfloat16, device="cuda")
    part = torch.zeros(64, dtype=torch.float64, device="cuda")
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    assert lib.vlbm_nn_loss(C.byref(g), yd.data_ptr(), td.data_ptr(), cd.data_ptr(), 12, float(inv), dz.data_ptr(),
                            part.data_ptr(), 64, out.data_ptr(), None) == 0
    torch.cuda.synchronize()
    l, dy = nn.loss(y, t, chw, 12, float(inv))
    assert float(out.item()) == pytest.approx(l, rel=1e-6)
    _bf16_close(_interior(dz), dy)


def test_gpu_adam_matches_oracle():
    torch, _native, lib = _setup()
    rng = np.random.default_rng(4)
    n = 1000
    p, m, v, gr = (rng.normal(size=n).astype(np.float32) for _ in range(4))
    v = np.abs(v)
    dev = [torch.from_numpy(a.copy()).cuda() for a in (p, m, v, gr)]
    b1, b2, eps, step = 0.9, 0.999, 1e-8, 3
    c1, c2 = 1 / (1 - b1 ** step), 1 / (1 - b2 ** step)
    assert lib.vlbm_nn_adam(n, *[d.data_ptr() for d in dev], 1e-3, b1, b2, eps, c1, c2, None) == 0
    torch.cuda.synchronize()
    wp, wm, wv = nn.adam(p.astype(np.float64), m.astype(np.float64), v.astype(np.float64), gr.astype(np.float64),
                         1e-3, b1, b2, eps, step)
    np.testing.assert_allclose(dev[0].cpu().numpy(), wp, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(dev[1].cpu().numpy(), wm, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(dev[2].cpu().numpy(), wv, rtol=1e-6, atol=1e-7)


def _batch_problem(torch, op, rng):
    B = op.batch
    x = np.zeros((B,) + op.shape + (32,))
    x[..., :op.n_in] = rng.normal(size=(B,) + op.shape + (op.n_in,))
    x = nn.bf16(x)
    t = np.zeros((B,) + op.shape + (16,))
    t[..., :op.n_out] = rng.normal(size=(B,) + op.shape + (op.n_out,))
    t[..., 15] = 1.0
    return x, t


def test_gpu_network_gradients_match_oracle():
    torch, _native, lib = _setup()
    from curved_lbm.surrogate.network import ConvOperator

    op = ConvOperator((5, 4, 6), batch=2, layers=4, n_in=6, n_out=5, seed=3, out_scale=1.0)
    rng = np.random.default_rng(5)
    x, t = _batch_problem(torch, op, rng)
    chw = np.zeros(32)
    chw[:5] = 1.0
    inv = 1.0 / float(x[..., 0].size)
    op.set_input(_padded(torch, x, torch.bfloat16))
    op.forward()
    op.loss_and_backward(_padded(torch, t, torch.float32), torch.from_numpy(chw).float().cuda(), inv)
    torch.cuda.synchronize()
    layers = [op.layer_weights(k) for k in range(op.layers)]
    y, acts, zs = nn.network_forward(x, layers, round_bf16=True)
    _, dy = nn.loss(y, t, chw, 5, inv)
    grads = nn.network_backward(dy, layers, acts, zs, round_bf16=True)
    got_y = _interior(op.y)[..., :5]
    assert np.linalg.norm(got_y - y[..., :5]) <= 2e-2 * np.linalg.norm(y[..., :5])
    for k in range(op.layers):
        gw = op.grads[k, :27 * 1024].cpu().numpy().reshape(27, 32, 32)
        gb = op.grads[k, 27 * 1024:].cpu().numpy()
        assert np.linalg.norm(gw - grads[k][0]) <= 3e-2 * np.linalg.norm(grads[k][0]), k
        assert np.linalg.norm(gb - grads[k][1]) <= 3e-2 * np.linalg.norm(grads[k][1]), k


def test_gpu_training_lowers_the_loss():
    torch, _native, lib = _setup()
    from curved_lbm.surrogate.network import ConvOperator

    op = ConvOperator((8, 8, 8), batch=4, layers=6, n_in=4, n_out=3, seed=1)
    rng = np.random.default_rng(6)
    B = 4
    X = np.stack(np.meshgrid(*(np.linspace(-1, 1, 8),) * 3, indexing="ij"), -1)
    x = np.zeros((B, 8, 8, 8, 32))
    amp = rng.uniform(0.5, 1.5, B)
    x[..., :3] = X
    x[..., 3] = amp[:, None, None, None]
    t = np.zeros((B, 8, 8, 8, 16))
    t[..., 0] = amp[:, None, None, None] * np.sin(2 * X[..., 0])
    t[..., 1] = X[..., 1] * X[..., 2]
    t[..., 2] = amp[:, None, None, None] ** 2 * X[..., 2]
    t[..., 15] = 1.0
    chw = torch.zeros(32, device="cuda")
    chw[:3] = 1.0
    op.set_input(_padded(torch, nn.bf16(x), torch.bfloat16))
    td = _padded(torch, t, torch.float32)
    inv = 1.0 / (B * 512 * 3)
    losses = []
    for it in range(300):
        op.forward()
        losses.append(op.loss_and_backward(td, chw, inv).item())
        op.adam_step(3e-3)
    assert np.isfinite(losses).all()
    assert losses[-1] < 0.05 * losses[0], (losses[0], losses[-1])
