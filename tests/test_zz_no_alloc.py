"""lt_observables_noalloc (the observables pass through a fixed device scratch array,
no allocation) against lt_observables, through the C ABI: equal bit for bit on two
calls in a row, the second on a smaller lattice (the array keeps the first call's
contents), and a lattice beyond the array's size is refused."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def flush(msg):
    print(msg, flush=True)


def observables_both(la, f):
    from lettuce_amd import _lib
    nt = f.shape[0]
    a = np.zeros(4 + 2 * nt)
    b = np.full(4 + 2 * nt, np.nan)
    _lib.call("lt_observables", f.handle, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    _lib.call("lt_observables_noalloc", f.handle, b.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return a, b


def test_noalloc_observables_equal_allocating_pass(gpu_ctx):
    import lettuce_amd as la
    for shape, seed in (((16, 16, 16, 16), 3), ((8, 6, 4, 10), 4)):
        flush(f"noalloc observables {shape}")
        f = la.GaugeField(gpu_ctx, shape)
        f.SetToHaar(seed)
        la.StoutSmearingN(f, f, 2, 0.1)
        a, b = observables_both(la, f)
        f.close()
        assert np.array_equal(a, b), np.max(np.abs(a - b))
        assert a[0] > 0.0 and np.all(np.isfinite(b))


def test_noalloc_observables_size_limit(gpu_ctx):
    """(76, 76, 76, 2) needs 31,848 doubles of scratch and fits; (74900, 2, 2, 2) needs
    14 * 74900 + 4 = 1,048,604 > 2^20 and is refused before any launch."""
    import lettuce_amd as la
    flush("noalloc observables: size limit")
    f = la.GaugeField(gpu_ctx, (76, 76, 76, 2))
    f.SetToHaar(5)
    a, b = observables_both(la, f)
    f.close()
    assert np.array_equal(a, b)
    g = la.GaugeField(gpu_ctx, (74900, 2, 2, 2))
    g.SetToIdentity()
    out = np.zeros(4 + 2 * 74900)
    from lettuce_amd import _lib
    with pytest.raises(la.LettuceError, match="scratch"):
        _lib.call("lt_observables_noalloc", g.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    g.close()
