"""The device's wave-level primitives (DPP reductions and the wave prefix
sum that the binned raster reserves its event slots with) against plain
sums over each 64-lane wave of one 256-thread workgroup."""
import ctypes as C

import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_wave_primitives(seed):
    hip = H.load_hip()
    rng = np.random.default_rng(seed)
    x = rng.integers(-100000, 100000, 256).astype(np.int32)
    if seed == 2:
        x[:] = rng.integers(0, 3, 256)  # small counts, as the raster's event counts
    out = np.zeros(6 * 256, dtype=np.int32)
    rc = hip.shape_hip_wave_selftest(x.ctypes.data_as(C.POINTER(C.c_int)), out.ctypes.data_as(C.POINTER(C.c_int)))
    assert rc == 0
    out = out.reshape(256, 6)
    w = x.reshape(4, 64)
    assert np.array_equal(out[:, 0], np.cumsum(w, axis=1).ravel())
    assert np.array_equal(out[:, 1], np.repeat(w.min(axis=1), 64))
    assert np.array_equal(out[:, 2], np.repeat(w.max(axis=1), 64))
    assert np.array_equal(out[:, 3], np.repeat(w.sum(axis=1), 64))
    f = w.astype(np.float32)
    assert np.array_equal(out[:, 4].view(np.float32), np.repeat(f.min(axis=1), 64))
    assert np.array_equal(out[:, 5].view(np.float32), np.repeat(f.max(axis=1), 64))
