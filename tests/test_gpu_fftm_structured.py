"""The reference's Fftm unit test restated on the HIP path: five structured
rows (zeros, a scaled impulse, a sum of decimated combs, a constant, one
sinusoid; tests/fftm.cpp:62-91) through forward Fftm along rows and along
columns, compared with the FP64 DFT at the reference's < -100 dB
criterion, and the inverse transform (scale 1/N) brought back to the input
(tests/fftm.cpp:111-200, 355-399).  The reference's sizes with a prime
factor above 5 (68, 252; 242 for the real transform) are outside this
backend's lengths: the plan reports them unsupported, which makes the plugin
decline them and OpenVSIP fall back to its next backend."""
import numpy as np
import pytest

import oracle as o

pytestmark = pytest.mark.gpu

DB = -100.0


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def structured_rows(n, scale=1.0, dtype=np.complex64):
    """setup_data_x (tests/fftm.cpp:62-91): 5 rows of length n."""
    x = np.zeros((5, n), dtype=np.complex128)
    x[1, 0] = scale
    x[2, 0] = 1.0
    x[2, 0:n:1] += 3.0
    for step, val in ((4, -2.0), (13, 7.0), (27, -15.0), (37, 31.0)):
        if n > step:
            x[2, 0:step * (n // step):step] += val
    x[3, :] = scale
    x[4, :] = np.sin(3.1415926535898 * np.arange(n) * 4 / n)
    return x.astype(dtype)


@pytest.mark.parametrize("n", [64, 256, 128, 512, 18])
@pytest.mark.parametrize("scale", [1.0, 2.5])
def test_fftm_structured_rows(ovx, n, scale):
    x = structured_rows(n, scale)
    fwd = ovx.Fftm(5, n, ovx.ROW, ovx.FFT_FWD, 1.0)
    inv = ovx.Fftm(5, n, ovx.ROW, ovx.FFT_INV, 1.0 / n)
    y = host(fwd(dev(x)))
    ref = o.fftm64(x.astype(np.complex128), o.FWD, 1.0)
    print("rows", n, scale, "db", o.diff_db(y, ref), flush=True)
    assert o.diff_db(y, ref) < DB
    back = host(inv(dev(y)))
    assert o.diff_db(back, x) < DB


@pytest.mark.parametrize("n", [256, 18])
def test_fftm_structured_columns(ovx, n):
    """test_by_ref_y: the same five signals as the columns of an n x 5
    matrix stored column-major, transformed along columns."""
    import torch
    sig = structured_rows(n)            # (5, n): signal k is column k
    fwd = ovx.Fftm(n, 5, ovx.COL, ovx.FFT_FWD, 1.0)
    inv = ovx.Fftm(n, 5, ovx.COL, ovx.FFT_INV, 1.0 / n)
    xcm = dev(sig).t()                  # (n, 5) view, column-major
    out = torch.empty((5, n), dtype=torch.complex64, device="cuda")
    fwd(xcm, out.t())
    y = host(out)                       # row k = transform of signal k
    ref = o.fftm64(sig.astype(np.complex128), o.FWD, 1.0)
    assert o.diff_db(y, ref) < DB
    back = torch.empty((5, n), dtype=torch.complex64, device="cuda")
    inv(out.t(), back.t())
    assert o.diff_db(host(back), sig) < DB


@pytest.mark.parametrize("n", [128, 16, 36])
def test_fftm_structured_real(ovx, n):
    """test_real: real rows to n/2 + 1 bins and back."""
    import torch
    x = structured_rows(n).real.astype(np.float32)
    fwd = ovx.FftmReal(5, n, ovx.ROW, ovx.FFT_FWD, 1.0)
    inv = ovx.FftmReal(5, n, ovx.ROW, ovx.FFT_INV, 1.0 / n)
    y = fwd(dev(x))
    ref = np.fft.rfft(x.astype(np.float64), axis=1)
    assert o.diff_db(host(y), ref) < DB
    back = host(inv(y))
    assert o.diff_db(back.astype(np.complex128), x.astype(np.complex128)) < DB


@pytest.mark.parametrize("n", [68, 252])
def test_fftm_lengths_with_large_prime_factors_are_declined(ovx, n):
    from openvsip_amd import OvxUnsupported
    with pytest.raises(OvxUnsupported):
        ovx.Fftm(5, n, ovx.ROW, ovx.FFT_FWD, 1.0)


def test_fftm_real_length_242_is_declined(ovx):
    from openvsip_amd import OvxUnsupported
    with pytest.raises(OvxUnsupported):
        ovx.FftmReal(5, 242, ovx.ROW, ovx.FFT_FWD, 1.0)
