"""The reference's boundary regressions on this path, restated on the HIP
backend: Fftm on sub-views whose rows start at an element offset inside a
wider matrix (row stride size + gap, first column `align`), out of place and
in place (tests/regressions/fftm_unaligned.cpp:20-118: 64 x 256 with (gap,
align) = (0, 0), (1, 1), (1, 0)), plus the same on the four-step and
hierarchical lengths; a fastconv whose input, output and spectrum views all
start at odd element offsets; the 1-D Fft on offset views and in place on a
matrix row, and Fftm in place on a sub-view (tests/regressions/
fft_unaligned.cpp, fft_ip_subview.cpp).  Elements outside the views stay as
they were."""
import numpy as np
import pytest

import oracle as o
from openvsip_amd import workloads as wl

pytestmark = pytest.mark.gpu

DB = -100.0


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


CASES = [(64, 256, 0, 0), (64, 256, 1, 1), (64, 256, 1, 0),
         (16, 16384, 3, 1), (4, 65536, 5, 3), (8, 6144, 2, 1)]


@pytest.mark.parametrize("rows,size,gap,align", CASES)
def test_fftm_unaligned_out_of_place(ovx, rows, size, gap, align):
    """test_fftm_op_align: row i of the input is the constant i + 1, so the
    transform is (i + 1) * size at frequency 0 and zero elsewhere; also the
    whole view against the FP64 oracle on random data."""
    import torch
    f = ovx.Fftm(rows, size, ovx.ROW, ovx.FFT_FWD, 1.0)
    inp = torch.zeros((rows, size + gap), dtype=torch.complex64, device="cuda")
    out = torch.full((rows, size + gap), complex(7, -7), dtype=torch.complex64, device="cuda")
    for i in range(rows):
        inp[i] = i + 1
    f(inp[:, align:align + size], out[:, align:align + size])
    y = host(out)
    for i in range(rows):
        assert y[i, align] == (i + 1) * size, (i, y[i, align])
    outside = np.ones(size + gap, bool)
    outside[align:align + size] = False
    assert np.all(y[:, outside] == complex(7, -7))
    # random data through the same views
    rng = np.random.default_rng(size + gap + align)
    x = wl.cgauss(rng, (rows, size + gap))
    inp.copy_(torch.from_numpy(x))
    f(inp[:, align:align + size], out[:, align:align + size])
    ref = o.fftm64(x[:, align:align + size].astype(np.complex128), o.FWD)
    assert o.diff_db(ref, host(out)[:, align:align + size]) < DB


@pytest.mark.parametrize("rows,size,gap,align", CASES)
def test_fftm_unaligned_in_place(ovx, rows, size, gap, align):
    """test_fftm_ip_align: a view of ones transforms in place to `size` at
    frequency 0; the columns outside the view keep their values."""
    import torch
    f = ovx.Fftm(rows, size, ovx.ROW, ovx.FFT_FWD, 1.0)
    m = torch.full((rows, size + gap), complex(-3, 5), dtype=torch.complex64, device="cuda")
    m[:, align:align + size] = 1
    f.in_place(m[:, align:align + size])
    y = host(m)
    assert np.all(y[:, align] == size)
    # the other bins are sums of rounded roots of unity times 1
    assert np.all(np.abs(y[:, align + 1:align + size]) < 1e-5 * size)
    outside = np.ones(size + gap, bool)
    outside[align:align + size] = False
    assert np.all(y[:, outside] == complex(-3, 5))


@pytest.mark.parametrize("n", [2048, 16384, 65536])
def test_fastconv_odd_offset_views(ovx, n):
    """Input, output and spectrum views that start one element into their
    allocations (8-byte but not 16-byte aligned) with row strides n + 3."""
    import torch
    rows = 6
    rng = np.random.default_rng(n)
    x = wl.cgauss(rng, (rows, n))
    w = wl.replica_spectrum(wl.lfm_chirp(n, n // 8), n)
    xb = torch.zeros((rows, n + 3), dtype=torch.complex64, device="cuda")
    yb = torch.full((rows, n + 3), complex(9, 9), dtype=torch.complex64, device="cuda")
    wb = torch.zeros(n + 1, dtype=torch.complex64, device="cuda")
    xb[:, 1:n + 1] = torch.from_numpy(x)
    wb[1:] = torch.from_numpy(w.astype(np.complex64))
    fc = ovx.Fastconv(rows, n)
    fc(xb[:, 1:n + 1], wb[1:], out=yb[:, 1:n + 1])
    y = host(yb)
    ref = o.fastconv64(x, w)
    assert o.rel_l2(y[:, 1:n + 1], ref) < 1e-5
    assert np.all(y[:, 0] == complex(9, 9)) and np.all(y[:, n + 1:] == complex(9, 9))


def _fft_in_place(f, t):
    import ctypes
    import torch
    from openvsip_amd import _lib
    _lib.check(_lib.lib().ovx_fft_in_place(f._h, ctypes.c_void_p(t.data_ptr()), t.stride(0), f.n,
                                           ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))


@pytest.mark.parametrize("align", [0, 1])
def test_fft_unaligned(ovx, align):
    """tests/regressions/fft_unaligned.cpp: a 256-point Fft of ones on a view
    starting `align` elements into its vector, out of place and in place;
    the transform is 256 at frequency 0."""
    import torch
    size = 256
    f = ovx.Fft(size, ovx.FFT_FWD, 1.0)
    inp = torch.ones(size + align, dtype=torch.complex64, device="cuda")
    out = torch.zeros(size + align, dtype=torch.complex64, device="cuda")
    f(inp[align:], out[align:])
    assert host(out)[align] == size
    _fft_in_place(f, inp[align:])
    y = host(inp)
    assert y[align] == size
    assert np.all(np.abs(y[align + 1:]) < 1e-5 * size)


def test_fft_in_place_on_vector_and_matrix_row(ovx):
    """tests/regressions/fft_ip_subview.cpp:17-47: the constant (1, -2) of
    length 64 transformed in place, as a vector and as the last row of a
    32 x 64 matrix: (64, -128) at frequency 0, zero elsewhere; the other
    rows of the matrix keep their values."""
    import torch
    rows, size = 32, 64
    f = ovx.Fft(size, ovx.FFT_FWD, 1.0)
    vec = torch.full((size,), complex(1, -2), dtype=torch.complex64, device="cuda")
    mat = torch.full((rows, size), complex(5, 5), dtype=torch.complex64, device="cuda")
    mat[rows - 1] = complex(1, -2)
    _fft_in_place(f, vec)
    _fft_in_place(f, mat[rows - 1])
    for y in (host(vec), host(mat)[rows - 1]):
        assert y[0] == complex(size, -2 * size)
        assert np.all(np.abs(y[1:]) < 1e-5 * size)
    assert np.all(host(mat)[: rows - 1] == complex(5, 5))


def test_fftm_in_place_on_subview(ovx):
    """tests/regressions/fft_ip_subview.cpp:50-96: Fftm (scale 1/cols) in
    place on a 16 x 64 matrix and on the same-sized sub-view [8:24, 32:96] of
    a 32 x 128 matrix, row r holding the constant r: frequency 0 of row r is
    exactly r in both, and the big matrix outside the view keeps -101."""
    import torch
    rows, cols = 16, 64
    f = ovx.Fftm(rows, cols, ovx.ROW, ovx.FFT_FWD, 1.0 / cols)
    inout = torch.zeros((rows, cols), dtype=torch.complex64, device="cuda")
    big = torch.full((2 * rows, 2 * cols), complex(-101, 0), dtype=torch.complex64, device="cuda")
    view = big[rows // 2:rows // 2 + rows, cols // 2:cols // 2 + cols]
    for r in range(rows):
        inout[r] = r
        view[r] = r
    f.in_place(inout)
    f.in_place(view)
    a, b = host(inout), host(big)
    v = b[rows // 2:rows // 2 + rows, cols // 2:cols // 2 + cols]
    for r in range(rows):
        assert a[r, 0] == r and v[r, 0] == r, r
    mask = np.ones(b.shape, bool)
    mask[rows // 2:rows // 2 + rows, cols // 2:cols // 2 + cols] = False
    assert np.all(b[mask] == complex(-101, 0))
