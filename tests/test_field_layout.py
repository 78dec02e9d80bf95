"""DeviceField presents the reference's (nx, ny, nz) indexing whatever the storage layout."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from waam_solver_amd.field import LAYOUT_XYZ, LAYOUT_ZYX, DeviceField, dims_of, layout_of


@pytest.mark.parametrize("layout", [LAYOUT_XYZ, LAYOUT_ZYX])
def test_round_trip_and_indexing(layout):
    rng = np.random.default_rng(3)
    a = rng.random((2, 3, 4, 5)).astype(np.float32)
    f = DeviceField.zeros(a.shape, torch.float32, "cpu", layout=layout)
    f.from_numpy(a)
    assert f.shape == a.shape and f.dims == (3, 4, 5) and layout_of(f) == layout
    assert f.tensor.shape == ((2, 5, 4, 3) if layout == LAYOUT_ZYX else a.shape)
    assert np.array_equal(f.to_numpy(), a)
    assert f[1, 2, 3, 4] == a[1, 2, 3, 4]
    for key in [(1,), (1, 2), (slice(None), 1), (0, slice(1, 3), 2)]:
        assert np.array_equal(f[key].numpy(), a[key]), key
    f[0, 1, 2, 3] = -1.0
    a[0, 1, 2, 3] = -1.0
    f[1, 2] = np.full((4, 5), 7.0, np.float32)
    a[1, 2] = 7.0
    assert np.array_equal(f.to_numpy(), a)


def test_zyx_storage_is_x_fastest():
    f = DeviceField.zeros((3, 4, 5), torch.float32, "cpu", layout=LAYOUT_ZYX)
    f[2, 1, 0] = 1.0
    flat = f.tensor.reshape(-1)
    assert int(torch.nonzero(flat)[0]) == (0 * 4 + 1) * 3 + 2  # ((k*ny + j)*nx + i)
    assert dims_of(f) == (3, 4, 5)
    assert dims_of(torch.zeros(19, 3, 4, 5)) == (3, 4, 5) and layout_of(torch.zeros(1)) == 0


def test_flags_read_back_as_int32():
    f = DeviceField.zeros((2, 2, 2), torch.uint8, "cpu", np_dtype=np.int32, layout=LAYOUT_ZYX)
    f[1, 0, 1] = 2
    a = f.to_numpy()
    assert a.dtype == np.int32 and a[1, 0, 1] == 2 and f[1, 0, 1] == 2
