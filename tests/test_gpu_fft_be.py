"""The reference's backend cross-check, tests/fft_be.cpp, restated on the HIP
entry points: ramp inputs (fft_be.cpp:167-183: element k of a vector is k,
row r of a matrix is r, r + 1, ...), 1-D transforms of length 16 and 8
(tests/fft_be.cpp:515-600: the length-8 case is a stride-2 subview, which
OpenVSIP's DDA hands a backend as dense data because query_layout requests
unit stride), Fftm of 8 x 16 along either axis (fft_be.cpp:925-960),
forward and inverse c->c by reference and in place, r->c and c->r, in
interleaved and split storage; the reference's criterion, error_db < -100
against the FP64 DFT."""
import ctypes

import numpy as np
import pytest

import oracle as o
from openvsip_amd import _lib

pytestmark = pytest.mark.gpu

DB = -100.0


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def P(t):
    return ctypes.c_void_p(t.data_ptr())


def ramp_matrix(rows, cols):
    return (np.arange(rows)[:, None] + np.arange(cols)[None, :]).astype(np.float64)


@pytest.mark.parametrize("n", [16, 8])
@pytest.mark.parametrize("inverse", [False, True])
@pytest.mark.parametrize("split", [False, True])
def test_fft_1d_c2c_by_ref_and_in_place(ovx, n, inverse, split):
    import torch
    lib = _lib.lib()
    x = np.arange(n).astype(np.complex64)
    ref = o.fftm64(x[None, :].astype(np.complex128), o.INV if inverse else o.FWD)[0]
    f = ovx.Fft(n, ovx.FFT_INV if inverse else ovx.FFT_FWD, 1.0)
    if split:
        xr, xi = torch.from_numpy(x.real.copy()).cuda(), torch.from_numpy(x.imag.copy()).cuda()
        yr, yi = torch.zeros_like(xr), torch.zeros_like(xi)
        _lib.check(lib.ovx_fft_split_out_of_place(f._h, P(xr), P(xi), 1, P(yr), P(yi), 1, n, None))
        assert o.diff_db(ref, host(yr) + 1j * host(yi)) < DB
        _lib.check(lib.ovx_fft_split_in_place(f._h, P(xr), P(xi), 1, n, None))
        assert o.diff_db(ref, host(xr) + 1j * host(xi)) < DB
    else:
        xd = torch.from_numpy(x).cuda()
        assert o.diff_db(ref, host(f(xd))) < DB
        _lib.check(lib.ovx_fft_in_place(f._h, P(xd), 1, n, None))
        assert o.diff_db(ref, host(xd)) < DB


@pytest.mark.parametrize("n", [16, 8])
@pytest.mark.parametrize("split", [False, True])
def test_fft_1d_real(ovx, n, split):
    """r->c of the ramp, and c->r of the ramp over n/2 + 1 values."""
    import torch
    lib = _lib.lib()
    h = n // 2 + 1
    xr = np.arange(n).astype(np.float32)
    fwd = ovx.FftmReal(1, n, ovx.ROW, ovx.FFT_FWD)
    inv = ovx.FftmReal(1, n, ovx.ROW, ovx.FFT_INV)
    xh = np.arange(h).astype(np.complex64)
    if split:
        sr, si = torch.zeros(h, device="cuda"), torch.zeros(h, device="cuda")
        _lib.check(lib.ovx_fftm_r2c_split(fwd._h, P(torch.from_numpy(xr).cuda()), n, 1, P(sr), P(si), h, 1, 1, n,
                                          None))
        spec = host(sr) + 1j * host(si)
        back = torch.zeros(n, device="cuda")
        hr = torch.from_numpy(xh.real.copy()).cuda()
        hi = torch.from_numpy(xh.imag.copy()).cuda()
        _lib.check(lib.ovx_fftm_c2r_split(inv._h, P(hr), P(hi), h, 1, P(back), n, 1, 1, n, None))
        y = host(back)
    else:
        spec = host(fwd(torch.from_numpy(xr[None, :].copy()).cuda()))[0]
        y = host(inv(torch.from_numpy(xh[None, :].copy()).cuda()))[0]
    assert o.diff_db(o.rfftm64(xr[None, :])[0], spec) < DB
    assert o.diff_db(o.irfftm64(xh[None, :], n)[0], y.astype(np.float64)) < DB


@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("inverse", [False, True])
@pytest.mark.parametrize("split", [False, True])
def test_fftm_8x16_c2c(ovx, axis, inverse, split):
    """Fftm of the 8 x 16 ramp along rows (axis 1, row-major data) or columns
    (axis 0: query_layout asks for column-major data, so the columns arrive
    contiguous), by reference and in place."""
    import torch
    lib = _lib.lib()
    rows, cols = 8, 16
    m = ramp_matrix(rows, cols).astype(np.complex64)
    d = o.INV if inverse else o.FWD
    direction = ovx.FFT_INV if inverse else ovx.FFT_FWD
    if axis == 1:
        lines, ref = m, o.fftm64(m.astype(np.complex128), d)
        f = ovx.Fftm(rows, cols, ovx.ROW, direction, 1.0)
        s0, s1 = cols, 1
    else:
        lines = np.ascontiguousarray(m.T)       # column c of m, contiguous
        ref = o.fftm64(lines.astype(np.complex128), d)
        f = ovx.Fftm(rows, cols, ovx.COL, direction, 1.0)
        s0, s1 = 1, rows
    if split:
        xr = torch.from_numpy(np.ascontiguousarray(lines.real)).cuda()
        xi = torch.from_numpy(np.ascontiguousarray(lines.imag)).cuda()
        yr, yi = torch.zeros_like(xr), torch.zeros_like(xi)
        _lib.check(lib.ovx_fftm_split_out_of_place(f._h, P(xr), P(xi), s0, s1, P(yr), P(yi), s0, s1, rows, cols,
                                                   None))
        assert o.diff_db(ref, host(yr) + 1j * host(yi)) < DB
        _lib.check(lib.ovx_fftm_split_in_place(f._h, P(xr), P(xi), s0, s1, rows, cols, None))
        assert o.diff_db(ref, host(xr) + 1j * host(xi)) < DB
    else:
        xd = torch.from_numpy(lines.copy()).cuda()
        yd = torch.zeros_like(xd)
        _lib.check(lib.ovx_fftm_out_of_place(f._h, P(xd), s0, s1, P(yd), s0, s1, rows, cols, None))
        assert o.diff_db(ref, host(yd)) < DB
        _lib.check(lib.ovx_fftm_in_place(f._h, P(xd), s0, s1, rows, cols, None))
        assert o.diff_db(ref, host(xd)) < DB


@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("split", [False, True])
def test_fftm_8x16_real(ovx, axis, split):
    """r->c and c->r Fftm of 8 x 16 along either axis (the complex side
    holds n/2 + 1 values per line)."""
    import torch
    lib = _lib.lib()
    rows, cols = 8, 16
    m = ramp_matrix(rows, cols).astype(np.float32)
    if axis == 1:
        lines, n, count = m, cols, rows
        pf = ovx.FftmReal(rows, cols, ovx.ROW, ovx.FFT_FWD)
        pi = ovx.FftmReal(rows, cols, ovx.ROW, ovx.FFT_INV)
    else:
        lines, n, count = np.ascontiguousarray(m.T), rows, cols
        pf = ovx.FftmReal(rows, cols, ovx.COL, ovx.FFT_FWD)
        pi = ovx.FftmReal(rows, cols, ovx.COL, ovx.FFT_INV)
    h = n // 2 + 1
    spec_ref = o.rfftm64(lines)
    xh = spec_ref.astype(np.complex64)          # a Hermitian-consistent spectrum
    back_ref = o.irfftm64(xh, n)
    # strides of each side: lines of n reals / h complex values, unit stride along a line
    rs = (n, 1) if axis == 1 else (1, n)
    cs = (h, 1) if axis == 1 else (1, h)
    xd = torch.from_numpy(lines.copy()).cuda()
    if split:
        sr, si = torch.zeros((count, h), device="cuda"), torch.zeros((count, h), device="cuda")
        _lib.check(lib.ovx_fftm_r2c_split(pf._h, P(xd), rs[0], rs[1], P(sr), P(si), cs[0], cs[1], rows, cols,
                                          None))
        spec = host(sr) + 1j * host(si)
        hr = torch.from_numpy(np.ascontiguousarray(xh.real)).cuda()
        hi = torch.from_numpy(np.ascontiguousarray(xh.imag)).cuda()
        back = torch.zeros((count, n), device="cuda")
        _lib.check(lib.ovx_fftm_c2r_split(pi._h, P(hr), P(hi), cs[0], cs[1], P(back), rs[0], rs[1], rows, cols,
                                          None))
    else:
        zd = torch.zeros((count, h), dtype=torch.complex64, device="cuda")
        _lib.check(lib.ovx_fftm_real_out_of_place(pf._h, P(xd), rs[0], rs[1], P(zd), cs[0], cs[1], rows, cols,
                                                  None))
        spec = host(zd)
        back = torch.zeros((count, n), device="cuda")
        _lib.check(lib.ovx_fftm_real_out_of_place(pi._h, P(torch.from_numpy(xh).cuda()), cs[0], cs[1], P(back),
                                                  rs[0], rs[1], rows, cols, None))
    assert o.diff_db(spec_ref, spec) < DB
    assert o.diff_db(back_ref, host(back).astype(np.float64)) < DB
