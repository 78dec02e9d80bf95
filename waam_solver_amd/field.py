"""Device fields: torch tensors on the ROCm device with the field API the reference's callers use.

The reference keeps every quantity in a Taichi field and its tests and tools touch fields
through ``to_numpy``, ``from_numpy``, ``f[i, j, k]``, ``buf[None]``, ``fill`` and ``copy_from``
(e.g. reference validation/test_bead_freeze.py:20-36, twin.py:721-746). ``DeviceField`` keeps
that surface over a contiguous torch tensor and exposes the raw device pointer the C ABI takes.
Cell flags are stored as uint8 and read back as int32, the reference's flag dtype.
"""

from __future__ import annotations

import numpy as np
import torch


class DeviceField:
    __slots__ = ("tensor", "_np_dtype")

    def __init__(self, tensor: torch.Tensor, np_dtype=None):
        if not tensor.is_contiguous():
            raise ValueError("DeviceField needs a contiguous tensor")
        self.tensor = tensor
        self._np_dtype = np_dtype

    # ---------------------------------------------------------------- construction
    @classmethod
    def zeros(cls, shape, dtype=torch.float32, device=None, np_dtype=None) -> "DeviceField":
        return cls(torch.zeros(shape, dtype=dtype, device=device), np_dtype)

    # ---------------------------------------------------------------- metadata
    @property
    def shape(self) -> tuple[int, ...]:
        return tuple(self.tensor.shape)

    @property
    def dtype(self):
        return self.tensor.dtype

    @property
    def ptr(self) -> int:
        return self.tensor.data_ptr()

    def __len__(self) -> int:
        return self.tensor.shape[0]

    # ---------------------------------------------------------------- host transfer
    def to_numpy(self) -> np.ndarray:
        a = self.tensor.detach().cpu().numpy()
        if self._np_dtype is not None:
            a = a.astype(self._np_dtype)
        return a

    def from_numpy(self, arr) -> None:
        a = np.asarray(arr)
        if tuple(a.shape) != self.shape:
            raise ValueError(f"shape {a.shape} does not match field shape {self.shape}")
        t = torch.from_numpy(np.ascontiguousarray(a)).to(self.tensor.dtype)
        self.tensor.copy_(t)

    def to_torch(self) -> torch.Tensor:
        return self.tensor

    # ---------------------------------------------------------------- element access
    def _key(self, key):
        if key is None:
            return ()
        return key

    def __getitem__(self, key):
        v = self.tensor[self._key(key)]
        if v.dim() == 0:
            x = v.item()
            return int(x) if self._np_dtype is not None and np.issubdtype(self._np_dtype, np.integer) else x
        return v

    def __setitem__(self, key, value) -> None:
        self.tensor[self._key(key)] = value

    def fill(self, value) -> None:
        self.tensor.fill_(value)

    def copy_from(self, other: "DeviceField") -> None:
        self.tensor.copy_(other.tensor)


def P(x) -> int:
    """Device pointer of a DeviceField or tensor (0 for None).

    Only contiguous tensors resident on the GPU are accepted: a host pointer handed to a kernel
    would be dereferenced on the device.
    """
    if x is None:
        return 0
    t = x.tensor if isinstance(x, DeviceField) else x
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"expected a DeviceField or tensor, got {type(x).__name__}")
    if t.device.type != "cuda":
        raise ValueError(f"kernel argument lives on {t.device}, not on the GPU")
    if not t.is_contiguous():
        raise ValueError("kernel arguments must be contiguous tensors")
    return t.data_ptr()


def tensor_of(x) -> torch.Tensor:
    return x.tensor if isinstance(x, DeviceField) else x
