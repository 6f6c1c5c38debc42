"""lbm_compute_residuals_noalloc against lbm_compute_residuals, through the
C ABI: the same values bit for bit (same kernels, the partial sums in the
library's fixed scratch instead of a per-call allocation) on two calls in a
row, the second on a smaller lattice, for both residual kinds; lattices above
the scratch's maximum are refused with an error and no device work."""
import ctypes as C

import numpy as np
import pytest

import parallellbmcranfield_amd as lbm
from hipmem import DeviceArray

pytestmark = pytest.mark.gpu


def _state(n, m, seed):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((n * m, 9))
    metaf = f + 1e-3 * rng.standard_normal((n * m, 9))
    bid = rng.integers(0, 4, n * m).astype(np.int32)
    return f, metaf, bid, rng.standard_normal(n * m), rng.standard_normal(n * m)


def _both(n, m, seed, kind):
    L = lbm.lib()
    dev = [DeviceArray(a) for a in _state(n, m, seed)]
    outs = []
    for fn in (L.lbm_compute_residuals, L.lbm_compute_residuals_noalloc):
        out = DeviceArray(nbytes=32)
        rc = fn(*(d.ptr for d in dev), n, m, 2, kind, out.ptr, None)
        assert rc == lbm.OK, lbm.lib().lbm_last_error()
        outs.append(out.get(np.float64, 4))
        out.free()
    for d in dev:
        d.free()
    return outs


@pytest.mark.parametrize("kind", [lbm.RESID_FULL, lbm.RESID_REFERENCE])
def test_noalloc_equals_the_allocating_entry_point(kind):
    print("noalloc kind", kind, flush=True)
    a1, b1 = _both(613, 407, 1, kind)     # 249 491 cells: 275 partial rows
    a2, b2 = _both(57, 31, 2, kind)       # a smaller lattice second: 2 rows
    assert np.array_equal(a1, b1) and np.array_equal(a2, b2)
    assert not np.array_equal(a1, a2)
    assert np.all(np.isfinite(a1)) and np.all(np.isfinite(a2))


def test_noalloc_refuses_lattices_above_its_scratch():
    L = lbm.lib()
    small = DeviceArray(nbytes=4096)
    out = DeviceArray(nbytes=32)
    n, m = 5462, 5461  # 29 828 982 cells > LBM_NOALLOC_MAX_CELLS
    rc = L.lbm_compute_residuals_noalloc(small.ptr, small.ptr, None, None, None, n, m, 0,
                                         lbm.RESID_FULL, out.ptr, None)
    assert rc == lbm.EINVAL
    assert b"exceed" in L.lbm_last_error()
    small.free()
    out.free()
