"""lh_plane_transform_noalloc (the plane transforms with their twiddle tables
in the library's fixed device array, no allocation) against the allocating
entry points lh_four_to_phys_planes / lh_phys_to_four_planes and the oracle:
two calls in a row per direction, the second on smaller planes, so the
shared array is rewritten between them."""
import ctypes

import numpy as np
import pytest

from oracle import pyoracle as po
from rnl_multiblock_nofft_amd import _lib

pytestmark = pytest.mark.gpu

SHAPES = [(32, 48, 48, 9), (16, 24, 24, 3)]     # (N1, Ngal1, Ngal2, planes); second smaller


def _run(fn, *args):
    _lib.check(fn(*args))


def _planes(shape, seed):
    N1, Ng1, Ng2, np_ = shape
    rng = np.random.default_rng(seed)
    return rng.standard_normal((np_, Ng2, Ng1 + 2))


class _Dev:
    """device buffer from the library's own C-ABI helpers (one HIP runtime
    in the process, whatever ran before)"""

    def __init__(self, arr):
        self.L = _lib.load()
        self.n = arr.nbytes
        self.p = ctypes.c_void_p()
        _run(self.L.lh_dev_alloc, self.n, ctypes.byref(self.p))
        _run(self.L.lh_copy_to_device, self.p, arr.ctypes.data_as(ctypes.c_void_p), self.n)

    def get(self, shape):
        out = np.empty(shape)
        _run(self.L.lh_copy_to_host, out.ctypes.data_as(ctypes.c_void_p), self.p, self.n)
        return out

    def free(self):
        _run(self.L.lh_dev_free, self.p)


@pytest.mark.parametrize("direction", [1, -1])
def test_noalloc_matches_allocating_transforms(direction):
    L = _lib.load()
    ref_fn = L.lh_four_to_phys_planes if direction > 0 else L.lh_phys_to_four_planes
    for q, shape in enumerate(SHAPES):
        N1, Ng1, Ng2, n = shape
        print("noalloc dir", direction, "shape", shape, flush=True)
        x = _planes(shape, 10 * q + (direction > 0))
        a, b = _Dev(x), _Dev(x)
        _run(ref_fn, N1, Ng1, Ng2, a.p, n, None)
        _run(L.lh_plane_transform_noalloc, direction, N1, Ng1, Ng2, b.p, n, None)
        ya, yb = a.get(x.shape), b.get(x.shape)     # synchronous copies
        a.free()
        b.free()
        e = np.abs(ya - yb).max() / np.abs(ya).max()
        print("  vs allocating entry point", e, flush=True)
        assert e <= 1e-14, (shape, e)
        # the first plane against the oracle's four_to_phys / phys_to_four
        o = po.Oracle(N1=N1, N2=2 * Ng2 // 3, ny=9, Re=100.0, Ngal1=Ng1, Ngal2=Ng2)
        want = o.four_to_phys(x[0]) if direction > 0 else o.phys_to_four(x[0])
        m = np.zeros((Ng2, Ng1 + 2), bool)
        m[:, :Ng1 if direction > 0 else N1 + 2] = True
        eo = np.abs(yb[0] - want)[m].max() / np.abs(want[m]).max()
        print("  vs oracle", eo, flush=True)
        assert eo <= 1e-13, (shape, eo)


def test_noalloc_rejects_sizes_beyond_its_array():
    L = _lib.load()
    a = _Dev(np.zeros(8))
    with pytest.raises(_lib.LhError):
        _run(L.lh_plane_transform_noalloc, 1, 2048, 3072, 48, a.p, 1, None)
    with pytest.raises(_lib.LhError):
        _run(L.lh_plane_transform_noalloc, 0, 32, 48, 48, a.p, 1, None)
    a.free()
