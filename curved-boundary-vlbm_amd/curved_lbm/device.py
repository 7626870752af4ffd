"""Device arrays of the solver: torch tensors on the HIP device, exposed with
the small array interface the reference solver attributes offer
(``.numpy()``, ``.shape``, ``.zero_()``, assignment from host arrays).

PyTorch only supplies device memory and the stream here; all arithmetic of
the time step runs in libvlbm.so.

Every array belongs to the stream that was current when it was created (a
solver creates its arrays on its own stream), and every torch operation the
array performs (assignment, zeroing, snapshots) is issued on that stream, so
it is ordered with the solver's library launches whichever stream the caller
has current.
"""

from __future__ import annotations

import numpy as np
import torch

_NP_TO_TORCH = {
    np.dtype(np.float64): torch.float64,
    np.dtype(np.float32): torch.float32,
    np.dtype(np.int32): torch.int32,
}


def hip_device(device: str | None) -> torch.device:
    dev = torch.device(device or "cuda:0")
    if dev.type != "cuda":
        raise ValueError(f"the solver runs on a HIP device (cuda:N), not {dev}")
    if not torch.cuda.is_available():
        raise RuntimeError("no HIP device is available; the solver has no CPU path")
    return dev


class DeviceArray:
    """A device tensor with numpy-style host access, bound to a stream."""

    def __init__(self, tensor: torch.Tensor, stream=None):
        self.tensor = tensor
        if stream is None and tensor.is_cuda:
            stream = torch.cuda.current_stream(tensor.device)
        self.stream = stream

    def _on_stream(self):
        return torch.cuda.stream(self.stream) if self.stream is not None else _NullContext()

    @classmethod
    def zeros(cls, shape, dtype, device) -> "DeviceArray":
        return cls(torch.zeros(shape, dtype=_NP_TO_TORCH[np.dtype(dtype)], device=device))

    @classmethod
    def from_numpy(cls, arr: np.ndarray, dtype, device) -> "DeviceArray":
        host = np.ascontiguousarray(np.asarray(arr, dtype=dtype))
        return cls(torch.from_numpy(host).to(device))

    def snapshot(self, out: torch.Tensor | None = None) -> torch.Tensor:
        """A device copy of the array taken on its stream (into ``out`` when
        given), ordered after the launches already queued there."""
        with self._on_stream():
            if out is None:
                return self.tensor.clone()
            out.copy_(self.tensor)
            return out

    @property
    def shape(self) -> tuple[int, ...]:
        return tuple(self.tensor.shape)

    @property
    def dtype(self):
        return self.tensor.dtype

    @property
    def ptr(self) -> int:
        return int(self.tensor.data_ptr()) if self.tensor.numel() else 0

    def numpy(self) -> np.ndarray:
        torch.cuda.synchronize(self.tensor.device)
        return self.tensor.detach().cpu().numpy()

    def assign(self, arr) -> None:
        if isinstance(arr, DeviceArray):
            arr = arr.tensor
        with self._on_stream():
            if isinstance(arr, torch.Tensor):
                src = arr.to(device=self.tensor.device, dtype=self.tensor.dtype)
            else:
                src = torch.from_numpy(np.ascontiguousarray(np.asarray(arr))).to(
                    device=self.tensor.device, dtype=self.tensor.dtype
                )
            if tuple(src.shape) != self.shape:
                raise ValueError(f"shape mismatch: {tuple(src.shape)} vs {self.shape}")
            self.tensor.copy_(src)

    def zero_(self) -> None:
        with self._on_stream():
            self.tensor.zero_()

    def fill_(self, value) -> None:
        with self._on_stream():
            self.tensor.fill_(value)

    def __repr__(self) -> str:
        return f"DeviceArray(shape={self.shape}, dtype={self.dtype}, device={self.tensor.device})"


class _NullContext:
    def __enter__(self):
        return None

    def __exit__(self, *exc):
        return False


class PopulationArray(DeviceArray):
    """Population planes as the reference sees them.  In FP32 mode the
    device holds float deviations from the rest state, and an owner whose
    ``populations_opaque`` is set (the 3D solver: adjacent planes paired per
    node) keeps its own layout in FP64 too; ``numpy()``/``assign`` then
    convert to and from the full double-precision planes through the HIP
    library."""

    def __init__(self, tensor: torch.Tensor, owner, stream=None):
        super().__init__(tensor, stream)
        self._owner = owner

    def _converted(self) -> bool:
        return self.tensor.dtype != torch.float64 or getattr(self._owner, "populations_opaque", False)

    def numpy(self) -> np.ndarray:
        if not self._converted():
            return super().numpy()
        return self._owner._populations_to_host(self)

    def assign(self, arr) -> None:
        if not self._converted():
            return super().assign(arr)
        if isinstance(arr, DeviceArray):
            arr = arr.numpy()
        if isinstance(arr, torch.Tensor):
            arr = arr.detach().to("cpu").numpy()
        self._owner._populations_from_host(self, np.asarray(arr, dtype=np.float64))


def copy(dst: DeviceArray, src) -> None:
    """dst <- src for device arrays, tensors or host arrays (like wp.copy)."""
    dst.assign(src)
