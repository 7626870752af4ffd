"""CPU FP64 restatement of the warm-start network kernels.

TEST INFRASTRUCTURE ONLY: the parity oracle of libvlbm's vlbm_nn_* entry
points (include/vlbm_nn.h); never imported by the product.  The network has
no counterpart in the reference (it starts every case from rest), so these
functions restate this build's own kernels from their definitions:

* conv(x, w, b): out[p, o] = b[o] + sum_{tap, i} w[tap, o, i] x[p + off(tap), i]
  over the 27 neighbour offsets, zero outside the grid (the zero halo);
* SiLU and its derivative;
* the weighted squared-error loss and its gradient;
* backward-data (transposed convolution times SiLU' of the layer below) and
  the weight / bias gradients;
* Adam.

Arrays are float64 in the grid layout (batch, nx, ny, nz, channels) without
the halo.  ``bf16`` rounds float values to the nearest BF16 (ties to even)
as the kernels do when they store a tensor, so that a kernel's output can be
compared with the oracle applied to the kernel's own inputs.  The backward
functions are pinned by central differences of ``loss`` through
``network_forward`` in pure FP64 (tests/test_oracle_nn.py).
"""

from __future__ import annotations

import numpy as np

TAPS = [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)]


def bf16(x: np.ndarray) -> np.ndarray:
    """Nearest BF16 value (ties to even) of each element, as float64."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    nan = np.isnan(f)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    out = (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    out[nan] = np.nan
    return out


def _shifted(x: np.ndarray, d) -> np.ndarray:
    """x[p + d] with zeros outside the grid (spatial axes 1..3)."""
    dx, dy, dz = d
    pad = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    X, Y, Z = x.shape[1:4]
    return pad[:, 1 + dx:1 + dx + X, 1 + dy:1 + dy + Y, 1 + dz:1 + dz + Z, :]


def conv(x: np.ndarray, w: np.ndarray, b: np.ndarray | None = None) -> np.ndarray:
    """3x3x3 convolution, w (27, out, in), tap order of TAPS."""
    out = np.zeros(x.shape[:4] + (w.shape[1],))
    for t, d in enumerate(TAPS):
        out += _shifted(x, d) @ w[t].T
    if b is not None:
        out += b
    return out


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def conv_backward_data(dz: np.ndarray, w: np.ndarray) -> np.ndarray:
    """d/dx of sum(dz * conv(x, w)): sum_tap w[tap]^T dz[p - off(tap)]."""
    out = np.zeros(dz.shape[:4] + (w.shape[2],))
    for t, d in enumerate(TAPS):
        out += _shifted(dz, (-d[0], -d[1], -d[2])) @ w[t]
    return out


def conv_backward_weight(x: np.ndarray, dz: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """(dw (27, out, in), db (out)) of sum(dz * conv(x, w, b))."""
    dw = np.zeros((27, dz.shape[-1], x.shape[-1]))
    flat_dz = dz.reshape(-1, dz.shape[-1])
    for t, d in enumerate(TAPS):
        dw[t] = flat_dz.T @ _shifted(x, d).reshape(-1, x.shape[-1])
    return dw, flat_dz.sum(axis=0)


def loss(y: np.ndarray, target: np.ndarray, chan_w: np.ndarray, n_out: int, inv_norm: float):
    """(loss, dloss/dy) of inv_norm * sum_p w_p sum_{c < n_out} chan_w[c] (y - t)^2,
    target (..., 16) with the voxel weight in channel 15."""
    wv = target[..., 15:16]
    d = y[..., :n_out] - target[..., :n_out]
    l = inv_norm * float(np.sum(wv * chan_w[:n_out] * d * d))
    g = np.zeros_like(y)
    g[..., :n_out] = 2.0 * wv * chan_w[:n_out] * d * inv_norm
    return l, g


def network_forward(x: np.ndarray, layers, *, round_bf16: bool = False):
    """Hidden layers z = conv(a) + b, a = SiLU(z); the last layer's output
    y = conv(a) + b (FP32 in the kernels).  Returns (y, [a_0 = x, a_1, ...],
    [z_1, ...]).  With round_bf16 every stored activation is rounded as the
    kernels store it."""
    r = bf16 if round_bf16 else (lambda v: v)
    acts, zs = [r(x)], []
    for k, (w, b) in enumerate(layers):
        wk = r(w)
        s = conv(acts[-1], wk, b)
        if k == len(layers) - 1:
            y = s.astype(np.float32).astype(np.float64) if round_bf16 else s
            return y, acts, zs
        zs.append(r(s))
        acts.append(r(silu(s)))
    raise ValueError("no layers")


def network_backward(dy: np.ndarray, layers, acts, zs, *, round_bf16: bool = False):
    """Gradients [(dw, db)] of every layer from dloss/dy, through stored
    activations; dz of hidden layers uses SiLU'(z) of the stored z."""
    r = bf16 if round_bf16 else (lambda v: v)
    grads = [None] * len(layers)
    dz = r(dy)
    for k in range(len(layers) - 1, -1, -1):
        grads[k] = conv_backward_weight(acts[k], dz)
        if k > 0:
            w = r(layers[k][0])
            dz = r(conv_backward_data(dz, w) * dsilu(zs[k - 1]))
    return grads


def adam(p, m, v, g, lr, b1, b2, eps, step):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1, c2 = 1.0 / (1.0 - b1 ** step), 1.0 / (1.0 - b2 ** step)
    return p - lr * (m * c1) / (np.sqrt(v * c2) + eps), m, v
