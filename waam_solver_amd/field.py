"""Device fields: torch tensors on the ROCm device with the field API the reference's callers use.

The reference keeps every quantity in a Taichi field and its tests and tools touch fields
through ``to_numpy``, ``from_numpy``, ``f[i, j, k]``, ``buf[None]``, ``fill`` and ``copy_from``
(e.g. reference validation/test_bead_freeze.py:20-36, twin.py:721-746). ``DeviceField`` keeps
that surface over a contiguous torch tensor and exposes the raw device pointer the C ABI takes.
Cell flags are stored as uint8 and read back as int32, the reference's flag dtype.

Grid fields carry a layout (the ABI's WAAM_LAYOUT_*): ``LAYOUT_XYZ`` stores (nx, ny, nz) with k
fastest, as the reference does; ``LAYOUT_ZYX`` stores (nz, ny, nx) with i fastest. Whatever the
storage, every accessor here speaks the reference's logical indexing: ``shape`` is
``(..., nx, ny, nz)``, ``to_numpy`` returns ``(..., nx, ny, nz)`` arrays and ``f[i, j, k]`` is cell
(i, j, k). Only the last three axes are spatial (populations are ``(19, nx, ny, nz)``).
"""

from __future__ import annotations

import numpy as np
import torch

LAYOUT_XYZ, LAYOUT_ZYX = 0, 1

# Count of host-initiated field writes (DeviceField setters and per-kernel library calls). The
# fused step keeps work lists across steps only while this count is unchanged.
_WRITES = [0]


def write_epoch() -> int:
    return _WRITES[0]


def note_write() -> None:
    _WRITES[0] += 1


def _spatial_perm(nd: int) -> tuple[int, ...]:
    """Axis permutation between logical (..., x, y, z) and stored (..., z, y, x) order (its own
    inverse)."""
    return tuple(range(nd - 3)) + (nd - 1, nd - 2, nd - 3)


class DeviceField:
    __slots__ = ("tensor", "_np_dtype", "layout")

    def __init__(self, tensor: torch.Tensor, np_dtype=None, layout: int | None = None):
        if not tensor.is_contiguous():
            raise ValueError("DeviceField needs a contiguous tensor")
        if layout not in (None, LAYOUT_XYZ, LAYOUT_ZYX):
            raise ValueError(f"unknown layout {layout}")
        if layout == LAYOUT_ZYX and tensor.dim() < 3:
            raise ValueError("a ZYX field needs three spatial axes")
        self.tensor = tensor
        self._np_dtype = np_dtype
        self.layout = layout

    # ---------------------------------------------------------------- construction
    @classmethod
    def zeros(cls, shape, dtype=torch.float32, device=None, np_dtype=None,
              layout: int | None = None) -> "DeviceField":
        """Zero field of logical shape `shape` (spatial axes last: nx, ny, nz)."""
        shape = tuple(shape)
        if layout == LAYOUT_ZYX:
            shape = tuple(shape[i] for i in _spatial_perm(len(shape)))
        return cls(torch.zeros(shape, dtype=dtype, device=device), np_dtype, layout)

    @property
    def _zyx(self) -> bool:
        return self.layout == LAYOUT_ZYX

    # ---------------------------------------------------------------- metadata
    @property
    def shape(self) -> tuple[int, ...]:
        s = tuple(self.tensor.shape)
        return tuple(s[i] for i in _spatial_perm(len(s))) if self._zyx else s

    @property
    def dims(self) -> tuple[int, int, int]:
        """(nx, ny, nz) of a grid field."""
        nx, ny, nz = self.shape[-3:]
        return nx, ny, nz

    @property
    def dtype(self):
        return self.tensor.dtype

    @property
    def ptr(self) -> int:
        return self.tensor.data_ptr()

    def __len__(self) -> int:
        return self.shape[0]

    # ---------------------------------------------------------------- host transfer
    def to_numpy(self) -> np.ndarray:
        a = self.tensor.detach().cpu().numpy()
        if self._zyx:
            a = np.ascontiguousarray(a.transpose(_spatial_perm(a.ndim)))
        if self._np_dtype is not None:
            a = a.astype(self._np_dtype)
        return a

    def from_numpy(self, arr) -> None:
        a = np.asarray(arr)
        if tuple(a.shape) != self.shape:
            raise ValueError(f"shape {a.shape} does not match field shape {self.shape}")
        if self._zyx:
            a = a.transpose(_spatial_perm(a.ndim))
        t = torch.from_numpy(np.ascontiguousarray(a)).to(self.tensor.dtype)
        self.tensor.copy_(t)
        note_write()

    def to_torch(self) -> torch.Tensor:
        """The stored tensor (for ZYX fields its axes are (..., z, y, x)). Writing through it is
        not seen by the fused step's cross-step bookkeeping: call ``WAAMTwin.sync()`` first."""
        note_write()
        return self.tensor

    # ---------------------------------------------------------------- element access
    def _view(self, key):
        """Stored-tensor view for a logical key, and the permutation that puts the view's axes
        back in logical order (None when there is nothing to permute)."""
        if key is None:
            key = ()
        if not isinstance(key, tuple):
            key = (key,)
        if not self._zyx:
            return self.tensor[key], None
        nd = self.tensor.dim()
        if any(k is Ellipsis or k is None for k in key) or len(key) > nd:
            raise IndexError("ZYX fields take plain integer/slice keys")
        full = list(key) + [slice(None)] * (nd - len(key))
        perm = _spatial_perm(nd)
        stored = [full[perm[a]] for a in range(nd)]
        view = self.tensor[tuple(stored)]
        # axes that survive, in stored order, as logical axis numbers
        kept = [perm[a] for a in range(nd) if not isinstance(stored[a], (int, np.integer))]
        order = sorted(range(len(kept)), key=lambda t: kept[t])
        return view, (tuple(order) if order != list(range(len(kept))) else None)

    def __getitem__(self, key):
        v, order = self._view(key)
        if order is not None:
            v = v.permute(order)
        if v.dim() == 0:
            x = v.item()
            return int(x) if self._np_dtype is not None and np.issubdtype(self._np_dtype, np.integer) else x
        return v

    def __setitem__(self, key, value) -> None:
        note_write()
        v, order = self._view(key)
        if order is not None:
            v = v.permute(order)
        if isinstance(value, (int, float, bool)):
            v.fill_(value)
        else:
            v.copy_(torch.as_tensor(np.asarray(value) if not isinstance(value, torch.Tensor)
                                    else value, device=v.device).to(v.dtype))

    def fill(self, value) -> None:
        note_write()
        self.tensor.fill_(value)

    def copy_from(self, other: "DeviceField") -> None:
        if other.layout != self.layout and self.tensor.dim() >= 3:
            raise ValueError("copy_from between fields of different layouts")
        note_write()
        self.tensor.copy_(other.tensor)


def P(x) -> int:
    """Device pointer of a DeviceField or tensor (0 for None).

    Only contiguous tensors resident on the GPU are accepted: a host pointer handed to a kernel
    would be dereferenced on the device. The address keeps no reference to the tensor: pass a
    tensor a name holds, never a temporary (``P(x.double())`` frees the copy, and the allocator
    may hand its memory to the next argument, before the call).
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


def layout_of(x) -> int:
    """Layout of a grid field (a bare tensor is taken to be in the reference layout)."""
    if isinstance(x, DeviceField) and x.layout is not None:
        return x.layout
    return LAYOUT_XYZ


def dims_of(x) -> tuple[int, int, int]:
    """Logical (nx, ny, nz) of a grid field or bare (nx, ny, nz)-shaped tensor."""
    if isinstance(x, DeviceField):
        return x.dims
    nx, ny, nz = tuple(x.shape)[-3:]
    return nx, ny, nz
