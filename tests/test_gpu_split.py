"""Split-complex storage (real and imaginary parts in two float arrays,
OpenVSIP's split_complex format) through the split entry points, which the
fftm_backend / fft_backend split overloads map to (backend.hpp:741-783; the
reference's FFTW backend runs split DFTs, fftw.cpp:945-1001).

Each split call is checked against the interleaved call of the same plan on
the same values -- equal bit for bit, since the split kernels run the same
arithmetic and only read and write the values as two float arrays (an exact
move) -- and against the FP64 oracle.  The 1-D
cases restate tests/fft.cpp:94-131 and 241-254 (sizes 64, 256, 128, 512 in
split storage; pattern `set 2` of setup_ptr, tests/fft.cpp:57-85; < -100 dB
against the FP64 DFT and after the inverse with scale 1/N)."""
import ctypes

import numpy as np
import pytest

import oracle as o
from openvsip_amd import _lib
from openvsip_amd import workloads as wl

pytestmark = pytest.mark.gpu

DB = -100.0
HOST = _lib.FLAG_HOST_POINTERS


def P(t):
    return ctypes.c_void_p(t.data_ptr())


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def split_dev(z):
    """(re, im) float32 device tensors of a complex array."""
    import torch
    z = np.ascontiguousarray(z, dtype=np.complex64)
    return (torch.from_numpy(np.ascontiguousarray(z.real)).cuda(),
            torch.from_numpy(np.ascontiguousarray(z.imag)).cuda())


def joined(re, im):
    return host(re).astype(np.float64) + 1j * host(im).astype(np.float64)


def as_c64(re, im):
    return (host(re) + 1j * host(im)).astype(np.complex64)


def plan_fftm(rows, cols, axis, direction, scale, flags=0):
    h = ctypes.c_void_p()
    _lib.check(_lib.lib().ovx_fftm_plan_create(ctypes.byref(h), rows, cols, axis, direction, scale, flags))
    return h


def setup_set2(n):
    """setup_ptr set 2 (tests/fft.cpp:67-76)."""
    x = np.zeros(n, np.complex128)
    x[0] = 1.0
    x[:] += 3.0
    for step, val in ((4, -2.0), (13, 7.0), (27, -15.0), (37, 31.0)):
        if n > step:
            x[0:step * (n // step):step] += val
    return x.astype(np.complex64)


@pytest.mark.parametrize("n", [64, 256, 128, 512, 6144, 16384, 65536])
def test_fft_split_by_reference(ovx, n):
    import torch
    lib = _lib.lib()
    x = setup_set2(n)
    f, i = ctypes.c_void_p(), ctypes.c_void_p()
    _lib.check(lib.ovx_fft_plan_create(ctypes.byref(f), n, _lib.FFT_FWD, 1.0, 0))
    _lib.check(lib.ovx_fft_plan_create(ctypes.byref(i), n, _lib.FFT_INV, 1.0 / n, 0))
    try:
        xr, xi = split_dev(x)
        yr, yi = torch.empty_like(xr), torch.empty_like(xi)
        _lib.check(lib.ovx_fft_split_out_of_place(f, P(xr), P(xi), 1, P(yr), P(yi), 1, n, None))
        ref = o.fftm64(x[None, :].astype(np.complex128), o.FWD)[0]
        assert o.diff_db(ref, joined(yr, yi)) < DB
        # the interleaved call of the same plan on the same values
        xc = torch.from_numpy(x).cuda()
        yc = torch.empty_like(xc)
        _lib.check(lib.ovx_fft_out_of_place(f, P(xc), 1, P(yc), 1, n, None))
        assert np.array_equal(as_c64(yr, yi), host(yc))
        # inverse in place brings the input back
        _lib.check(lib.ovx_fft_split_in_place(i, P(yr), P(yi), 1, n, None))
        assert o.diff_db(x.astype(np.complex128), joined(yr, yi)) < DB
    finally:
        lib.ovx_plan_destroy(f)
        lib.ovx_plan_destroy(i)


# rows, n, row gap: register Stockham, hierarchical and four-step paths (the
# four-step split passes run 32 lines per group: 256 x 256, 256 x 512 and
# 512 x 512 matrices, the last in 1024-thread groups)
CASES = [(2, 64, 0), (3, 512, 5), (7, 2048, 1), (16, 16384, 3), (4, 65536, 2), (5, 6144, 0), (3, 131072, 0),
         (2, 262144, 1)]


@pytest.mark.parametrize("rows,n,gap", CASES)
def test_fftm_split_rows_equal_interleaved(ovx, rows, n, gap):
    """Rows at stride n + gap, out of place and in place; the gap columns
    keep their values."""
    import torch
    lib = _lib.lib()
    ld = n + gap
    rng = np.random.default_rng(rows * n + gap)
    z = wl.cgauss(rng, (rows, ld))
    h = plan_fftm(rows, n, _lib.ROW, _lib.FFT_FWD, 1.0)
    try:
        xr, xi = split_dev(z)
        yr = torch.full((rows, ld), 4.0, device="cuda")
        yi = torch.full((rows, ld), -4.0, device="cuda")
        _lib.check(lib.ovx_fftm_split_out_of_place(h, P(xr), P(xi), ld, 1, P(yr), P(yi), ld, 1, rows, n, None))
        zc = torch.from_numpy(z).cuda()
        yc = torch.empty_like(zc)
        _lib.check(lib.ovx_fftm_out_of_place(h, P(zc), ld, 1, P(yc), ld, 1, rows, n, None))
        ys = as_c64(yr, yi)
        assert np.array_equal(ys[:, :n], host(yc)[:, :n])
        assert np.all(host(yr)[:, n:] == 4.0) and np.all(host(yi)[:, n:] == -4.0)
        ref = o.fftm64(z[:, :n].astype(np.complex128), o.FWD)
        assert o.rel_l2(ys[:, :n], ref) < 1e-5
        # in place on the split arrays
        _lib.check(lib.ovx_fftm_split_in_place(h, P(xr), P(xi), ld, 1, rows, n, None))
        assert np.array_equal(as_c64(xr, xi)[:, :n], ys[:, :n])
        assert np.array_equal(as_c64(xr, xi)[:, n:], z[:, n:])
    finally:
        lib.ovx_plan_destroy(h)


@pytest.mark.parametrize("n,cols", [(256, 5), (16384, 3)])
def test_fftm_split_columns_inverse(ovx, n, cols):
    """Column Fftm (unit stride down a column, column stride ld) in the
    inverse direction with scale 1/n."""
    import torch
    lib = _lib.lib()
    ld = n + 1
    rng = np.random.default_rng(n + cols)
    zt = wl.cgauss(rng, (cols, ld))      # column c of the n x cols matrix is zt[c, :n]
    h = plan_fftm(n, cols, _lib.COL, _lib.FFT_INV, 1.0 / n)
    try:
        xr, xi = split_dev(zt)
        yr, yi = torch.zeros_like(xr), torch.zeros_like(xi)
        _lib.check(lib.ovx_fftm_split_out_of_place(h, P(xr), P(xi), 1, ld, P(yr), P(yi), 1, ld, n, cols, None))
        zc = torch.from_numpy(zt).cuda()
        yc = torch.zeros_like(zc)
        _lib.check(lib.ovx_fftm_out_of_place(h, P(zc), 1, ld, P(yc), 1, ld, n, cols, None))
        assert np.array_equal(as_c64(yr, yi)[:, :n], host(yc)[:, :n])
        ref = o.fftm64(zt[:, :n].astype(np.complex128), o.INV, 1.0 / n)
        assert o.rel_l2(as_c64(yr, yi)[:, :n], ref) < 1e-5
    finally:
        lib.ovx_plan_destroy(h)


@pytest.mark.parametrize("n", [2048, 16384, 65536, 131072, 262144])
@pytest.mark.parametrize("matrix", [False, True])
def test_fastconv_split_equals_interleaved(ovx, n, matrix):
    import torch
    lib = _lib.lib()
    rows = 6
    rng = np.random.default_rng(n + matrix)
    x = wl.cgauss(rng, (rows, n + 2))
    w = wl.replica_spectrum(wl.lfm_chirp(n, n // 8), n).astype(np.complex64)
    wm = np.stack([np.roll(w, 5 * k) for k in range(rows)]).astype(np.complex64) if matrix else w
    ldw = n if matrix else 0
    h = ctypes.c_void_p()
    _lib.check(lib.ovx_fastconv_plan_create(ctypes.byref(h), rows, n, 1.0 / n, 0))
    try:
        xr, xi = split_dev(x)
        wr, wi = split_dev(wm)
        yr, yi = torch.zeros((rows, n + 1), device="cuda"), torch.zeros((rows, n + 1), device="cuda")
        _lib.check(lib.ovx_fastconv_split_exec(h, P(xr), P(xi), n + 2, P(wr), P(wi), ldw, P(yr), P(yi), n + 1,
                                               rows, None))
        xc, wc = torch.from_numpy(x).cuda(), torch.from_numpy(wm).cuda()
        yc = torch.zeros((rows, n + 1), dtype=torch.complex64, device="cuda")
        _lib.check(lib.ovx_fastconv_exec(h, P(xc), n + 2, P(wc), ldw, P(yc), n + 1, rows, None))
        ys = as_c64(yr, yi)
        assert np.array_equal(ys[:, :n], host(yc)[:, :n])
        assert np.all(ys[:, n:] == 0)
        if matrix:
            ref = np.stack([o.fastconv64(x[k:k + 1, :n], wm[k])[0] for k in range(rows)])
        else:
            ref = o.fastconv64(x[:, :n], w)
        assert o.rel_l2(ys[:, :n], ref) < 1e-5
    finally:
        lib.ovx_plan_destroy(h)


def test_split_host_pointers(ovx):
    """Host (numpy) split arrays with padded strides, as a split-storage
    VSIPL++ block hands them over through DDA: Fftm in place and fastconv;
    the padding keeps its values."""
    lib = _lib.lib()
    rows, n, ld = 4, 1024, 1030
    rng = np.random.default_rng(77)
    z = wl.cgauss(rng, (rows, ld))
    re, im = np.ascontiguousarray(z.real), np.ascontiguousarray(z.imag)
    h = plan_fftm(rows, n, _lib.ROW, _lib.FFT_FWD, 1.0, HOST)
    try:
        _lib.check(lib.ovx_fftm_split_in_place(h, re.ctypes.data, im.ctypes.data, ld, 1, rows, n, None))
        ref = o.fftm64(z[:, :n].astype(np.complex128), o.FWD)
        assert o.rel_l2((re + 1j * im)[:, :n], ref) < 1e-5
        assert np.array_equal(re[:, n:], z.real[:, n:]) and np.array_equal(im[:, n:], z.imag[:, n:])
    finally:
        lib.ovx_plan_destroy(h)
    m = 16384
    x = wl.cgauss(rng, (rows, m))
    w = wl.replica_spectrum(wl.lfm_chirp(m, m // 8), m).astype(np.complex64)
    xr, xi = np.ascontiguousarray(x.real), np.ascontiguousarray(x.imag)
    wr, wi = np.ascontiguousarray(w.real), np.ascontiguousarray(w.imag)
    yr, yi = np.full((rows, m + 3), 9.0, np.float32), np.full((rows, m + 3), 9.0, np.float32)
    fc = ctypes.c_void_p()
    _lib.check(lib.ovx_fastconv_plan_create(ctypes.byref(fc), rows, m, 1.0 / m, HOST))
    try:
        _lib.check(lib.ovx_fastconv_split_exec(fc, xr.ctypes.data, xi.ctypes.data, m, wr.ctypes.data,
                                               wi.ctypes.data, 0, yr.ctypes.data, yi.ctypes.data, m + 3, rows,
                                               None))
        assert o.rel_l2((yr + 1j * yi)[:, :m], o.fastconv64(x, w)) < 1e-5
        assert np.all(yr[:, m:] == 9.0) and np.all(yi[:, m:] == 9.0)
    finally:
        lib.ovx_plan_destroy(fc)


def test_split_refusals(ovx):
    import torch
    lib = _lib.lib()
    n = 16384
    h = ctypes.c_void_p()
    _lib.check(lib.ovx_fastconv_plan_create(ctypes.byref(h), 2, n, 1.0 / n, _lib.FLAG_BF16X3))
    try:
        t = torch.zeros(2 * n, device="cuda")
        assert lib.ovx_fastconv_split_exec(h, P(t), P(t), n, P(t), P(t), 0, P(t), P(t), n, 2,
                                           None) == _lib.OVX_ERR_UNSUPPORTED
    finally:
        lib.ovx_plan_destroy(h)
    f = plan_fftm(2, 64, _lib.ROW, _lib.FFT_FWD, 1.0)
    try:
        assert lib.ovx_fftm_split_out_of_place(f, None, None, 64, 1, None, None, 64, 1, 2, 64,
                                               None) == _lib.OVX_ERR_INVALID
        # an Fftm plan is not an Fft plan
        t = torch.zeros(64, device="cuda")
        assert lib.ovx_fft_split_in_place(f, P(t), P(t), 1, 64, None) == _lib.OVX_ERR_INVALID
    finally:
        lib.ovx_plan_destroy(f)


def plan_real(rows, n, direction, scale, flags=0):
    h = ctypes.c_void_p()
    _lib.check(_lib.lib().ovx_fftm_real_plan_create(ctypes.byref(h), rows, n, _lib.ROW, direction, scale, flags))
    return h


# real lengths through the register-Stockham, hierarchical and four-step
# paths (half lengths 8 ... 65536); 128 and 16 are tests/fft.cpp:256-257
@pytest.mark.parametrize("rows,n", [(1, 128), (3, 16), (5, 2048), (4, 32768), (2, 131072)])
def test_real_fftm_split_equals_interleaved(ovx, rows, n):
    """r2c into split arrays and c2r from them, with padded complex strides,
    against the interleaved real Fftm of the same plans: equal bit for bit;
    the round trip returns the input."""
    import torch
    lib = _lib.lib()
    h = n // 2 + 1
    ldc = h + 3
    rng = np.random.default_rng(n + rows)
    x = rng.standard_normal((rows, n)).astype(np.float32)
    fwd = plan_real(rows, n, _lib.FFT_FWD, 1.0)
    inv = plan_real(rows, n, _lib.FFT_INV, 1.0 / n)
    try:
        xd = torch.from_numpy(x).cuda()
        sr = torch.full((rows, ldc), 2.0, device="cuda")
        si = torch.full((rows, ldc), -2.0, device="cuda")
        _lib.check(lib.ovx_fftm_r2c_split(fwd, P(xd), n, 1, P(sr), P(si), ldc, 1, rows, n, None))
        zc = torch.zeros((rows, ldc), dtype=torch.complex64, device="cuda")
        _lib.check(lib.ovx_fftm_real_out_of_place(fwd, P(xd), n, 1, P(zc), ldc, 1, rows, n, None))
        spec = as_c64(sr, si)
        assert np.array_equal(spec[:, :h], host(zc)[:, :h])
        assert np.all(host(sr)[:, h:] == 2.0) and np.all(host(si)[:, h:] == -2.0)
        assert o.rel_l2(spec[:, :h], o.rfftm64(x)) < 1e-5
        back = torch.zeros((rows, n + 2), device="cuda")
        _lib.check(lib.ovx_fftm_c2r_split(inv, P(sr), P(si), ldc, 1, P(back), n + 2, 1, rows, n, None))
        ref_back = torch.zeros((rows, n + 2), device="cuda")
        _lib.check(lib.ovx_fftm_real_out_of_place(inv, P(zc), ldc, 1, P(ref_back), n + 2, 1, rows, n, None))
        assert np.array_equal(host(back), host(ref_back))
        assert o.rel_l2(host(back)[:, :n].astype(np.complex128), x.astype(np.complex128)) < 1e-5
        # the directions are the plans'
        assert lib.ovx_fftm_r2c_split(inv, P(xd), n, 1, P(sr), P(si), ldc, 1, rows, n,
                                      None) == _lib.OVX_ERR_INVALID
        assert lib.ovx_fftm_c2r_split(fwd, P(sr), P(si), ldc, 1, P(back), n + 2, 1, rows, n,
                                      None) == _lib.OVX_ERR_INVALID
    finally:
        lib.ovx_plan_destroy(fwd)
        lib.ovx_plan_destroy(inv)


def test_real_fftm_split_host_pointers(ovx):
    lib = _lib.lib()
    rows, n = 3, 4096
    h = n // 2 + 1
    rng = np.random.default_rng(5)
    x = np.full((rows, n + 4), 6.0, np.float32)
    x[:, :n] = rng.standard_normal((rows, n))
    fwd = plan_real(rows, n, _lib.FFT_FWD, 1.0, HOST)
    inv = plan_real(rows, n, _lib.FFT_INV, 1.0 / n, HOST)
    try:
        sr = np.full((rows, h + 1), 8.0, np.float32)
        si = np.full((rows, h + 1), 8.0, np.float32)
        _lib.check(lib.ovx_fftm_r2c_split(fwd, x.ctypes.data, n + 4, 1, sr.ctypes.data, si.ctypes.data, h + 1, 1,
                                          rows, n, None))
        assert o.rel_l2((sr + 1j * si)[:, :h], o.rfftm64(x[:, :n])) < 1e-5
        assert np.all(sr[:, h:] == 8.0) and np.all(si[:, h:] == 8.0)
        back = np.full((rows, n + 2), -3.0, np.float32)
        _lib.check(lib.ovx_fftm_c2r_split(inv, sr.ctypes.data, si.ctypes.data, h + 1, 1, back.ctypes.data, n + 2, 1,
                                          rows, n, None))
        assert o.rel_l2(back[:, :n].astype(np.complex128), x[:, :n].astype(np.complex128)) < 1e-5
        assert np.all(back[:, n:] == -3.0)
    finally:
        lib.ovx_plan_destroy(fwd)
        lib.ovx_plan_destroy(inv)
