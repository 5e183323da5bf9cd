"""hs_grid_slope_noalloc: the operator on caller-owned device buffers with no
device allocation (its descriptor and power map in the library's fixed
scratch array) equals hs_grid_slope bit for bit, on two calls in a row, the
second on a smaller model, and on a heterogeneous (microchannel) model; a
model whose vector does not fit the scratch, or whose device state is not
built, is refused with an error."""
import ctypes as C
import tempfile

import numpy as np
import pytest

from hotspot_amd import cases, iface
from helpers import bind_extensions, check_error, make, ptr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib(prod, gpu):
    L = bind_extensions(prod)
    L.lib.hs_grid_slope_noalloc.argtypes = [iface.RCModelP, C.c_void_p, C.c_void_p, C.c_void_p]
    L.lib.hs_grid_slope_noalloc.restype = C.c_int
    return L


@pytest.fixture(scope="module")
def hip():
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


class DevBuf:
    """A device buffer of float64 owned by the test (the caller)."""

    def __init__(self, hip, n):
        self.hip, self.n = hip, n
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(n, 1) * 8) == 0

    def put(self, a):
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, 1) == 0   # host to device

    def get(self):
        a = np.zeros(self.n)
        assert self.hip.hipMemcpy(a.ctypes.data, self.p, a.nbytes, 2) == 0   # device to host
        return a

    def free(self):
        self.hip.hipFree(self.p)


def slope_both(lib, hip, run, P, T):
    n = lib.lib.hs_grid_size(run.model)
    ref = np.zeros(n)
    lib.lib.hs_grid_slope(run.model, ptr(P), ptr(T), ptr(ref))   # also builds the device state
    check_error(lib)
    bp, bt, bs = DevBuf(hip, P.size), DevBuf(hip, n), DevBuf(hip, n)
    bp.put(P)
    bt.put(T)
    bs.put(np.full(n, np.nan))
    rc = lib.lib.hs_grid_slope_noalloc(run.model, bp.p, bt.p, bs.p)
    check_error(lib)
    assert rc == 0
    out = bs.get()
    for b in (bp, bt, bs):
        b.free()
    return out, ref


def test_noalloc_equals_grid_slope_twice(lib, hip):
    big, pb = cases.stacked_case(128, seed=3, n_dies=2, blocks_per_die=16)
    small = cases.default_stack_case(32)
    mc, pm = cases.microchannel_case(16, 5, tempfile.mkdtemp())
    for case, pw in ((big, pb), (small, [cases.rng(1).uniform(0.5, 5, 16)]), (mc, pm)):
        run = make(lib, case)
        P = run.power_vector(pw)
        n = lib.lib.hs_grid_size(run.model)
        T = 318.0 + 40.0 * cases.rng(9).random(n)
        out, ref = slope_both(lib, hip, run, P, T)
        assert np.array_equal(out, ref), case.name
        run.free()


def test_noalloc_refuses(lib, hip):
    # device state not built yet
    run = make(lib, cases.default_stack_case(32))
    n = lib.lib.hs_grid_size(run.model)
    b = DevBuf(hip, n)
    assert lib.lib.hs_grid_slope_noalloc(run.model, b.p, b.p, b.p) == -1
    with pytest.raises(RuntimeError, match="no device state"):
        check_error(lib)
    run.free()
    # 1024 x 1024 x 4 nodes: the power map does not fit the 32 MiB scratch
    run = make(lib, cases.default_stack_case(1024), tmp=None)
    assert lib.lib.hs_grid_slope_noalloc(run.model, b.p, b.p, b.p) == -1
    with pytest.raises(RuntimeError, match="bytes of scratch"):
        check_error(lib)
    run.free()
    b.free()
