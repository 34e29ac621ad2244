"""ovx_storage: a block's data in host memory and in HBM with the reference
storage_manager's rules (src/ovxx/storage/manager.hpp:207-360): a location is
refreshed only when it is stale, a writable pointer makes the other location
stale, write-only access copies nothing, and released user storage is
admitted back with or without an update.  The copies are counted, so each
test checks both the data (fastconv results against the FP64 oracle) and
that no copy happened that the rules do not call for."""
import ctypes

import numpy as np
import pytest

import oracle as o
from openvsip_amd import _lib
from openvsip_amd import workloads as wl
from openvsip_amd.api import Storage

pytestmark = pytest.mark.gpu

HOST, DEV = 1 << _lib.LOC_HOST, 1 << _lib.LOC_DEVICE
R, W, RW = _lib.ACCESS_READ, _lib.ACCESS_WRITE, _lib.ACCESS_READ_WRITE


def host_view(p, shape):
    count = int(np.prod(shape))
    f = np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_float)), shape=(2 * count,))
    return f.view(np.complex64).reshape(shape)


def fastconv_plan(rows, n):
    h = ctypes.c_void_p()
    _lib.check(_lib.lib().ovx_fastconv_plan_create(ctypes.byref(h), rows, n, 1.0 / n, 0))
    return h


def test_storage_moves_data_only_when_a_location_is_stale(ovx):
    lib = _lib.lib()
    rows, n = 8, 4096
    rng = np.random.default_rng(11)
    x = wl.cgauss(rng, (rows, n))
    w = wl.replica_spectrum(wl.lfm_chirp(n, n // 8), n).astype(np.complex64)
    nb = x.nbytes
    sx, sy, sw = Storage(nb, x), Storage(nb), Storage(w.nbytes, w)
    assert sx.info() == (HOST, 0, 0)
    assert sy.info() == (HOST | DEV, 0, 0)          # fresh: nothing to copy
    plan = fastconv_plan(rows, n)
    try:
        def run():
            dx, dw = sx.ptr(_lib.LOC_DEVICE, R), sw.ptr(_lib.LOC_DEVICE, R)
            dy = sy.ptr(_lib.LOC_DEVICE, W)
            _lib.check(lib.ovx_fastconv_exec(plan, dx, n, dw, 0, dy, n, rows, None))

        run()
        assert sx.info() == (HOST | DEV, nb, 0)      # one H2D of x
        assert sy.info() == (DEV, 0, 0)              # written on the device only
        run()                                        # x resident: no second copy
        assert sx.info() == (HOST | DEV, nb, 0)
        y = host_view(sy.ptr(_lib.LOC_HOST, R), (rows, n)).copy()
        assert sy.info() == (HOST | DEV, 0, nb)      # one D2H of y
        assert o.rel_l2(y, o.fastconv64(x, w)) < 1e-5
        host_view(sy.ptr(_lib.LOC_HOST, R), (rows, n))
        assert sy.info()[2] == nb                    # host valid: no copy
        # the host writes x: the device copy goes stale and is refreshed once
        hx = sx.ptr(_lib.LOC_HOST, RW)
        assert hx.value == x.ctypes.data and sx.info()[0] == HOST
        x *= 2
        run()
        run()
        assert sx.info() == (HOST | DEV, 2 * nb, 0)
        y2 = host_view(sy.ptr(_lib.LOC_HOST, R), (rows, n))
        assert o.rel_l2(y2, o.fastconv64(x, w)) < 1e-5
    finally:
        lib.ovx_plan_destroy(plan)


def test_storage_release_and_admit(ovx):
    lib = _lib.lib()
    n = 2048
    x = wl.cgauss(np.random.default_rng(12), (1, n))
    s = Storage(x.nbytes, x)
    d = s.ptr(_lib.LOC_DEVICE, RW)                   # device copy, host stale
    assert s.info() == (DEV, x.nbytes, 0)
    # run an in-place forward Fftm on the device copy
    plan = ctypes.c_void_p()
    _lib.check(lib.ovx_fftm_plan_create(ctypes.byref(plan), 1, n, _lib.ROW, _lib.FFT_FWD, 1.0, 0))
    try:
        _lib.check(lib.ovx_fftm_in_place(plan, d, n, 1, 1, n, None))
        ref = o.fftm64(x.astype(np.complex128), o.FWD)
        _lib.check(lib.ovx_storage_release(s._h, None))   # the array is brought up to date
        assert s.info() == (HOST | DEV, x.nbytes, x.nbytes)
        assert o.rel_l2(x, ref) < 1e-5
        p = ctypes.c_void_p()
        assert lib.ovx_storage_ptr(s._h, _lib.LOC_HOST, R, ctypes.byref(p), None) == _lib.OVX_ERR_STATE
        # admitted without update: both copies stay valid, nothing moves
        _lib.check(lib.ovx_storage_admit(s._h, 0))
        s.ptr(_lib.LOC_DEVICE, R)
        assert s.info() == (HOST | DEV, x.nbytes, x.nbytes)
        # released, changed, admitted with update: the device copy is refreshed
        _lib.check(lib.ovx_storage_release(s._h, None))
        x[:] = 1
        _lib.check(lib.ovx_storage_admit(s._h, 1))
        assert s.info()[0] == HOST
        d = s.ptr(_lib.LOC_DEVICE, RW)
        _lib.check(lib.ovx_fftm_in_place(plan, d, n, 1, 1, n, None))
        y = host_view(s.ptr(_lib.LOC_HOST, R), (1, n))
        assert y[0, 0] == n and np.all(np.abs(y[0, 1:]) < 1e-5 * n)
        assert s.info() == (HOST | DEV, 2 * x.nbytes, 2 * x.nbytes)
    finally:
        lib.ovx_plan_destroy(plan)


def test_storage_on_streams(ovx):
    """Copies run on the caller's stream: a copy queued on one stream is
    waited for by work on another that uses the same device copy."""
    import torch
    lib = _lib.lib()
    rows, n = 16, 16384
    rng = np.random.default_rng(13)
    x = wl.cgauss(rng, (rows, n))
    w = wl.replica_spectrum(wl.lfm_chirp(n, n // 8), n).astype(np.complex64)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    h1, h2 = ctypes.c_void_p(s1.cuda_stream), ctypes.c_void_p(s2.cuda_stream)
    sx, sw, sy = Storage(x.nbytes, x), Storage(w.nbytes, w), Storage(x.nbytes)
    plan = fastconv_plan(rows, n)
    try:
        sx.ptr(_lib.LOC_DEVICE, R, h1)               # H2D queued on s1
        sw.ptr(_lib.LOC_DEVICE, R, h1)
        dx, dw = sx.ptr(_lib.LOC_DEVICE, R, h2), sw.ptr(_lib.LOC_DEVICE, R, h2)  # s2 waits for it
        dy = sy.ptr(_lib.LOC_DEVICE, W, h2)
        _lib.check(lib.ovx_fastconv_exec(plan, dx, n, dw, 0, dy, n, rows, h2))
        y = host_view(sy.ptr(_lib.LOC_HOST, R, h2), (rows, n))
        assert o.rel_l2(y, o.fastconv64(x, w)) < 1e-5
        assert sx.info() == (HOST | DEV, x.nbytes, 0)
    finally:
        lib.ovx_plan_destroy(plan)


def test_storage_library_owned_host_memory(ovx):
    """Without a user array the host location is pinned memory the library
    allocates on first use; data written there survives a round trip through
    the device bit for bit."""
    n = 3000
    s = Storage(n * 8)
    h = host_view(s.ptr(_lib.LOC_HOST, W), (n,))
    v = wl.cgauss(np.random.default_rng(14), n)
    h[:] = v
    assert s.info() == (HOST, 0, 0)
    s.ptr(_lib.LOC_DEVICE, RW)                       # H2D; host stale
    back = host_view(s.ptr(_lib.LOC_HOST, R), (n,))  # D2H
    assert np.array_equal(back, v)
    assert s.info() == (HOST | DEV, n * 8, n * 8)


def test_storage_argument_errors(ovx):
    lib = _lib.lib()
    s = Storage(64)
    p = ctypes.c_void_p()
    assert lib.ovx_storage_ptr(s._h, 2, R, ctypes.byref(p), None) == _lib.OVX_ERR_INVALID
    assert lib.ovx_storage_ptr(s._h, _lib.LOC_HOST, 0, ctypes.byref(p), None) == _lib.OVX_ERR_INVALID
    assert lib.ovx_storage_release(s._h, None) == _lib.OVX_ERR_STATE   # no user array
    empty = Storage(0)
    _lib.check(lib.ovx_storage_ptr(empty._h, _lib.LOC_DEVICE, RW, ctypes.byref(p), None))
    assert p.value is None
