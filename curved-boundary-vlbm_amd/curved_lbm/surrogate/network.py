"""The convolutional neural operator on the HIP kernels of include/vlbm_nn.h.

Six 3x3x3 convolution layers of 32 channels with SiLU between them, on a
cell lattice (batch, nx, ny, nz) with a zero halo: the input carries up to
32 feature channels, the output layer's first ``n_out`` channels are the
prediction (FP32).  Weights are kept as FP32 masters for Adam and rounded to
BF16 for the MFMA kernels after every step.  Torch only holds the device
buffers; every layer, gradient, loss and update runs in libvlbm.so.
"""

from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from .. import _native

CH = 32
TAPS = 27
NW = TAPS * CH * CH
TARGET_CH = 16
WGRAD_PARTIAL = NW + CH


class ConvOperator:
    """Parameters, activations and gradients of the network for one grid
    shape and batch size (``resize_batch`` changes the batch)."""

    def __init__(self, shape, *, batch: int = 1, layers: int = 6, n_in: int = 12, n_out: int = 12,
                 device: str = "cuda:0", seed: int = 0, out_scale: float = 0.1):
        if layers < 2:
            raise ValueError("the network needs at least two layers")
        if not 1 <= n_in <= CH or not 1 <= n_out <= TARGET_CH - 1:
            raise ValueError("n_in must be in [1, 32] and n_out in [1, 15]")
        self.lib = _native.load_library()
        self.device = torch.device(device)
        self.shape = tuple(int(s) for s in shape)
        self.layers, self.n_in, self.n_out = int(layers), int(n_in), int(n_out)
        rng = np.random.default_rng(seed)
        master = np.zeros((self.layers, NW + CH), np.float32)
        for k in range(self.layers):
            fan_in = TAPS * (self.n_in if k == 0 else CH)
            w = rng.normal(0.0, math.sqrt(2.0 / fan_in), size=(TAPS, CH, CH))
            if k == 0:
                w[:, :, self.n_in:] = 0.0
            if k == self.layers - 1:
                w *= out_scale
            master[k, :NW] = w.reshape(-1)
        self.params = torch.from_numpy(master).to(self.device)  # (layers, NW + 32): weights then bias
        self.grads = torch.zeros_like(self.params)
        self.adam_m = torch.zeros_like(self.params)
        self.adam_v = torch.zeros_like(self.params)
        self.adam_t = 0
        self.w = torch.zeros((self.layers, NW), dtype=torch.bfloat16, device=self.device)
        self.wt = torch.zeros_like(self.w)
        self.stream = torch.cuda.current_stream(self.device)
        self._pack()
        self.batch = 0
        self.resize_batch(batch)
        self._loss = torch.zeros(1, dtype=torch.float64, device=self.device)
        self._loss_part = torch.zeros(1024, dtype=torch.float64, device=self.device)
        self._wpart = torch.zeros((256, WGRAD_PARTIAL), dtype=torch.float32, device=self.device)

    # ---- buffers -----------------------------------------------------------
    @property
    def padded_shape(self):
        return tuple(s + 2 for s in self.shape)

    def resize_batch(self, batch: int) -> None:
        batch = int(batch)
        if batch == self.batch:
            return
        self.batch = batch
        full = (batch,) + self.padded_shape
        bf = dict(dtype=torch.bfloat16, device=self.device)
        self.acts = [torch.zeros(full + (CH,), **bf) for _ in range(self.layers)]  # a_0 = input
        self.zs = [torch.zeros(full + (CH,), **bf) for _ in range(self.layers - 1)]
        self.dz = [torch.zeros(full + (CH,), **bf) for _ in range(2)]
        self.y = torch.zeros(full + (CH,), dtype=torch.float32, device=self.device)
        self.grid = _native.NNGrid(batch, *self.shape)

    def _st(self):
        return C.c_void_p(self.stream.cuda_stream or None)

    def _check(self, rc, what):
        _native.check(rc, what)

    def _pack(self) -> None:
        for k in range(self.layers):
            self._check(self.lib.vlbm_nn_pack_weights(self.params[k].data_ptr(), self.w[k].data_ptr(),
                                                      self.wt[k].data_ptr(), self._st()), "pack_weights")

    def _bias(self, k):
        return self.params[k, NW:].data_ptr()

    # ---- forward / backward --------------------------------------------------
    def set_input(self, x: torch.Tensor) -> None:
        """x: (batch, nx+2, ny+2, nz+2, 32) BF16 with a zero halo."""
        if tuple(x.shape) != tuple(self.acts[0].shape):
            raise ValueError(f"input must have shape {tuple(self.acts[0].shape)}")
        self.acts[0].copy_(x)

    def forward(self) -> torch.Tensor:
        g = C.byref(self.grid)
        for k in range(self.layers - 1):
            self._check(self.lib.vlbm_nn_conv_forward(g, self.acts[k].data_ptr(), self.w[k].data_ptr(), self._bias(k),
                                                      0, self.zs[k].data_ptr(), self.acts[k + 1].data_ptr(), None,
                                                      self._st()), "conv_forward")
        k = self.layers - 1
        self._check(self.lib.vlbm_nn_conv_forward(g, self.acts[k].data_ptr(), self.w[k].data_ptr(), self._bias(k), 1,
                                                  None, None, self.y.data_ptr(), self._st()), "conv_forward")
        return self.y

    def loss_and_backward(self, target: torch.Tensor, chan_w: torch.Tensor, inv_norm: float) -> torch.Tensor:
        """Loss of the last forward against target (batch, padded, 16) FP32
        (channel 15 = voxel weight); fills self.grads.  Returns the loss as
        a one-element device tensor (no host synchronisation)."""
        g = C.byref(self.grid)
        npart = int(self._loss_part.numel())
        self._check(self.lib.vlbm_nn_loss(g, self.y.data_ptr(), target.data_ptr(), chan_w.data_ptr(), self.n_out,
                                          float(inv_norm), self.dz[0].data_ptr(), self._loss_part.data_ptr(), npart,
                                          self._loss.data_ptr(), self._st()), "loss")
        cur = 0
        for k in range(self.layers - 1, -1, -1):
            gk = self.grads[k]
            self._check(self.lib.vlbm_nn_conv_backward_weight(
                g, self.acts[k].data_ptr(), self.dz[cur].data_ptr(), self._wpart.data_ptr(), int(self._wpart.shape[0]),
                gk.data_ptr(), gk[NW:].data_ptr(), self._st()), "conv_backward_weight")
            if k > 0:
                self._check(self.lib.vlbm_nn_conv_backward_data(
                    g, self.dz[cur].data_ptr(), self.wt[k].data_ptr(), self.zs[k - 1].data_ptr(),
                    self.dz[1 - cur].data_ptr(), self._st()), "conv_backward_data")
                cur = 1 - cur
        return self._loss

    def adam_step(self, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8) -> None:
        self.adam_t += 1
        c1, c2 = 1.0 / (1.0 - b1 ** self.adam_t), 1.0 / (1.0 - b2 ** self.adam_t)
        n = int(self.params.numel())
        self._check(self.lib.vlbm_nn_adam(n, self.params.data_ptr(), self.adam_m.data_ptr(), self.adam_v.data_ptr(),
                                          self.grads.data_ptr(), float(lr), float(b1), float(b2), float(eps),
                                          float(c1), float(c2), self._st()), "adam")
        self._pack()

    # ---- parameters ----------------------------------------------------------
    def state(self) -> dict:
        return {"params": self.params.detach().cpu().numpy(), "shape": np.array(self.shape),
                "layers": self.layers, "n_in": self.n_in, "n_out": self.n_out}

    def load_state(self, params: np.ndarray) -> None:
        p = torch.from_numpy(np.ascontiguousarray(params, dtype=np.float32)).to(self.device)
        if tuple(p.shape) != tuple(self.params.shape):
            raise ValueError(f"parameters must have shape {tuple(self.params.shape)}")
        self.params.copy_(p)
        self._pack()

    def layer_weights(self, k: int) -> tuple[np.ndarray, np.ndarray]:
        """FP32 master weights (27, out, in) and bias (32) of layer k."""
        p = self.params[k].detach().cpu().numpy()
        return p[:NW].reshape(TAPS, CH, CH).copy(), p[NW:].copy()
