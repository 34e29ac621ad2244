"""Plans created with OVX_FLAG_HOST_POINTERS take host (numpy) buffers: the
call stages them to the device, runs the same kernels and copies back before
it returns.  That is how the VSIPL++ plugin calls the backend (OpenVSIP's DDA
hands it host pointers), so every entry point the plugin uses is checked
here against the FP64 oracle, with strided rows and in place where the
entry point allows them."""
import ctypes

import numpy as np
import pytest

import oracle as o
from openvsip_amd import _lib
from openvsip_amd import workloads as wl

pytestmark = pytest.mark.gpu

HOST = _lib.FLAG_HOST_POINTERS
REL_L2 = 1e-5


def P(a):
    return ctypes.c_void_p(a.ctypes.data)


class Plan:
    def __init__(self):
        self.h = ctypes.c_void_p()

    def __del__(self):
        if self.h:
            _lib.lib().ovx_plan_destroy(self.h)


@pytest.mark.parametrize("n,ld", [(1024, 1030), (16384, 16384), (65536, 65540)])
def test_host_fftm_strided_and_in_place(ovx, n, ld):
    lib = _lib.lib()
    rows = 5
    rng = np.random.default_rng(n)
    big = wl.cgauss(rng, (rows, ld))
    p = Plan()
    _lib.check(lib.ovx_fftm_plan_create(ctypes.byref(p.h), rows, n, _lib.ROW, _lib.FFT_FWD, 1.0, HOST))
    out = np.full((rows, ld), 3 + 3j, np.complex64)
    _lib.check(lib.ovx_fftm_out_of_place(p.h, P(big), ld, 1, P(out), ld, 1, rows, n, None))
    ref = o.fftm64(big[:, :n].astype(np.complex128), o.FWD)
    assert o.rel_l2(out[:, :n], ref) < REL_L2
    assert np.all(out[:, n:] == 3 + 3j)
    inout = big.copy()
    _lib.check(lib.ovx_fftm_in_place(p.h, P(inout), ld, 1, rows, n, None))
    assert o.rel_l2(inout[:, :n], ref) < REL_L2
    assert np.array_equal(inout[:, n:], big[:, n:])


def test_host_fft_1d(ovx):
    lib = _lib.lib()
    n = 6144
    x = wl.cgauss(np.random.default_rng(1), n)
    p = Plan()
    _lib.check(lib.ovx_fft_plan_create(ctypes.byref(p.h), n, _lib.FFT_INV, 1.0 / n, HOST))
    y = np.empty(n, np.complex64)
    _lib.check(lib.ovx_fft_out_of_place(p.h, P(x), 1, P(y), 1, n, None))
    assert o.rel_l2(y, o.fftm64(x[None, :].astype(np.complex128), o.INV, 1.0 / n)[0]) < REL_L2


@pytest.mark.parametrize("n", [2048, 16384, 65536])
def test_host_fastconv_vector_and_matrix_spectrum(ovx, n):
    lib = _lib.lib()
    rows = 4
    rng = np.random.default_rng(n + 1)
    x = wl.cgauss(rng, (rows, n))
    w = wl.replica_spectrum(wl.lfm_chirp(n, n // 8), n).astype(np.complex64)
    p = Plan()
    _lib.check(lib.ovx_fastconv_plan_create(ctypes.byref(p.h), rows, n, 1.0 / n, HOST))
    y = np.empty_like(x)
    _lib.check(lib.ovx_fastconv_exec(p.h, P(x), n, P(w), 0, P(y), n, rows, None))
    assert o.rel_l2(y, o.fastconv64(x, w)) < REL_L2
    # coefficient matrix (one spectrum per row), in place
    wm = np.stack([np.roll(w, k) for k in range(rows)]).astype(np.complex64)
    z = x.copy()
    _lib.check(lib.ovx_fastconv_exec(p.h, P(z), n, P(wm), n, P(z), n, rows, None))
    ref = np.stack([o.fastconv64(x[k:k + 1], wm[k])[0] for k in range(rows)])
    assert o.rel_l2(z, ref) < REL_L2


@pytest.mark.parametrize("n", [2048, 65536])
def test_host_fastconv_strided_rows_keep_gaps(ovx, n):
    """Input, spectrum-matrix and output rows with padded strides: only the
    n elements of each row are read and written; the padding of the output
    keeps the caller's values."""
    lib = _lib.lib()
    rows = 3
    rng = np.random.default_rng(n + 2)
    xb = wl.cgauss(rng, (rows, n + 5))
    w = wl.replica_spectrum(wl.lfm_chirp(n, n // 8), n).astype(np.complex64)
    wb = np.zeros((rows, n + 7), np.complex64)
    for k in range(rows):
        wb[k, :n] = np.roll(w, 3 * k)
    yb = np.full((rows, n + 3), -4 + 2j, np.complex64)
    p = Plan()
    _lib.check(lib.ovx_fastconv_plan_create(ctypes.byref(p.h), rows, n, 1.0 / n, HOST))
    _lib.check(lib.ovx_fastconv_exec(p.h, P(xb), n + 5, P(wb), n + 7, P(yb), n + 3, rows, None))
    ref = np.stack([o.fastconv64(xb[k:k + 1, :n], wb[k, :n])[0] for k in range(rows)])
    assert o.rel_l2(yb[:, :n], ref) < REL_L2
    assert np.all(yb[:, n:] == -4 + 2j)


@pytest.mark.parametrize("n", [1024, 32768])
def test_host_real_fftm_round_trip(ovx, n):
    lib = _lib.lib()
    rows = 3
    x = np.random.default_rng(n).standard_normal((rows, n)).astype(np.float32)
    fwd, inv = Plan(), Plan()
    _lib.check(lib.ovx_fftm_real_plan_create(ctypes.byref(fwd.h), rows, n, _lib.ROW, _lib.FFT_FWD, 1.0,
                                             HOST))
    _lib.check(lib.ovx_fftm_real_plan_create(ctypes.byref(inv.h), rows, n, _lib.ROW, _lib.FFT_INV,
                                             1.0 / n, HOST))
    h = n // 2 + 1
    spec = np.empty((rows, h), np.complex64)
    _lib.check(lib.ovx_fftm_real_out_of_place(fwd.h, P(x), n, 1, P(spec), h, 1, rows, n, None))
    assert o.rel_l2(spec, o.rfftm64(x)) < REL_L2
    back = np.empty_like(x)
    _lib.check(lib.ovx_fftm_real_out_of_place(inv.h, P(spec), h, 1, P(back), n, 1, rows, n, None))
    assert o.rel_l2(back.astype(np.complex128), x.astype(np.complex128)) < REL_L2
    # padded line strides on both sides; the padding keeps its values
    xp = np.full((rows, n + 4), 7.0, np.float32)
    xp[:, :n] = x
    sp = np.full((rows, h + 2), 5 - 5j, np.complex64)
    _lib.check(lib.ovx_fftm_real_out_of_place(fwd.h, P(xp), n + 4, 1, P(sp), h + 2, 1, rows, n, None))
    assert o.rel_l2(sp[:, :h], o.rfftm64(x)) < REL_L2
    assert np.all(sp[:, h:] == 5 - 5j)
    bp = np.full((rows, n + 6), -1.0, np.float32)
    _lib.check(lib.ovx_fftm_real_out_of_place(inv.h, P(sp), h + 2, 1, P(bp), n + 6, 1, rows, n, None))
    assert o.rel_l2(bp[:, :n].astype(np.complex128), x.astype(np.complex128)) < REL_L2
    assert np.all(bp[:, n:] == -1.0)


@pytest.mark.parametrize("support", ["full", "same", "min"])
def test_host_correlation_and_convolution(ovx, support):
    lib = _lib.lib()
    sup = _lib.SUPPORT[support]
    rng = np.random.default_rng(len(support))
    m, n = 33, 5000
    ref, x = wl.cgauss(rng, m), wl.cgauss(rng, n)
    c = Plan()
    _lib.check(lib.ovx_corr_plan_create(ctypes.byref(c.h), sup, m, n, HOST))
    out = np.empty(lib.ovx_corr_output_size(sup, m, n), np.complex64)
    _lib.check(lib.ovx_corr_exec(c.h, _lib.BIAS["unbiased"], P(ref), P(x), P(out), None))
    assert o.rel_l2(out, o.corr64("unbiased", support, ref, x)) < REL_L2
    v = Plan()
    _lib.check(lib.ovx_conv_plan_create(ctypes.byref(v.h), _lib.SYMMETRY["nonsym"], sup, P(ref), m, n, 2, HOST))
    cout = np.empty(lib.ovx_conv_output_size(sup, m, n, 2), np.complex64)
    _lib.check(lib.ovx_conv_exec(v.h, P(x), P(cout), None))
    assert o.rel_l2(cout, o.conv64(support, o.conv_kernel("nonsym", ref), x, 2)) < REL_L2
