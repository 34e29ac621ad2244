"""GPU parity tests of the real <-> complex Fftm (r2c / c2r through a
half-length complex transform) against the FP64 oracle, which follows the
reference's FFTW r2c/c2r calls (src/ovxx/fftw/fftw.cpp:778-903) and is pinned
against numpy.fft.rfft / irfft and the reference's own r2c / c2r golden files
in tests/test_oracle.py.  Tolerance: rel-L2 <= 1e-5 (complex<float> bar); the
golden-file tests use the reference's < -100 dB criterion."""
import numpy as np
import pytest

import oracle as o

pytestmark = pytest.mark.gpu

REL_L2 = 1e-5


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def say(*a):
    print(*a, flush=True)


# n: real length; the complex half-length h = n/2 covers the register-Stockham
# (32, 6144), hierarchical (2048, 32768) and four-step (131072) paths
@pytest.mark.parametrize("n,rows", [(4, 3), (18, 7), (32, 5), (72, 3), (1024, 33), (2048, 16), (6144, 9),
                                    (32768, 12), (131072, 4)])
def test_r2c_rows_vs_fp64(ovx, n, rows):
    import torch
    rng = np.random.default_rng(n + rows)
    x = rng.standard_normal((rows, n)).astype(np.float32)
    f = ovx.FftmReal(rows, n, ovx.ROW, ovx.FFT_FWD)
    y = host(f(torch.from_numpy(x).cuda()))
    ref = o.rfftm64(x)
    err = o.rel_l2(y, ref)
    say(f"r2c n={n} rows={rows} plan={f.name} rel_l2={err:.3e}")
    assert y.shape == (rows, n // 2 + 1)
    assert err < REL_L2


@pytest.mark.parametrize("n,rows", [(4, 3), (18, 7), (32, 5), (72, 3), (1024, 33), (2048, 16), (6144, 9),
                                    (32768, 12), (131072, 4)])
def test_c2r_rows_vs_fp64(ovx, n, rows):
    import torch
    rng = np.random.default_rng(2 * n + rows)
    xh = (rng.standard_normal((rows, n // 2 + 1)) + 1j * rng.standard_normal((rows, n // 2 + 1))).astype(np.complex64)
    f = ovx.FftmReal(rows, n, ovx.ROW, ovx.FFT_INV, scale=1.0 / n)
    y = host(f(torch.from_numpy(xh).cuda()))
    ref = o.irfftm64(xh, n, scale=1.0 / n)
    err = o.rel_l2(y, ref)
    say(f"c2r n={n} rows={rows} plan={f.name} rel_l2={err:.3e}")
    assert y.shape == (rows, n)
    assert y.dtype == np.float32
    assert err < REL_L2


def test_real_round_trip_and_scale(ovx):
    import torch
    rows, n = 7, 4096
    rng = np.random.default_rng(4)
    x = rng.standard_normal((rows, n)).astype(np.float32)
    fwd = ovx.FftmReal(rows, n, ovx.ROW, ovx.FFT_FWD, scale=2.0)
    inv = ovx.FftmReal(rows, n, ovx.ROW, ovx.FFT_INV, scale=0.5 / n)
    back = host(inv(fwd(torch.from_numpy(x).cuda())))
    assert o.rel_l2(back, x) < REL_L2


def test_real_column_axis_and_strides(ovx):
    import torch
    rng = np.random.default_rng(5)
    n, cols = 512, 6
    # column lines are contiguous: the rows x cols real matrix is stored
    # column-major
    xc = rng.standard_normal((cols, n)).astype(np.float32)
    xd = torch.from_numpy(xc).cuda().t()          # shape (n, cols), strides (1, n)
    f = ovx.FftmReal(n, cols, ovx.COL, ovx.FFT_FWD)
    y = f(xd)                                     # (n/2+1, cols) column-major
    assert tuple(y.shape) == (n // 2 + 1, cols)
    assert o.rel_l2(host(y).T, o.rfftm64(xc)) < REL_L2
    # rows with a padded (even) stride on both sides
    rows = 5
    xp = np.zeros((rows, n + 6), np.float32)
    xp[:, :n] = rng.standard_normal((rows, n))
    yp = torch.zeros((rows, n // 2 + 1 + 3), dtype=torch.complex64, device="cuda")
    g = ovx.FftmReal(rows, n, ovx.ROW, ovx.FFT_FWD)
    g(torch.from_numpy(xp).cuda()[:, :n], out=yp[:, : n // 2 + 1])
    yh = host(yp)
    assert o.rel_l2(yh[:, : n // 2 + 1], o.rfftm64(xp[:, :n])) < REL_L2
    assert np.all(yh[:, n // 2 + 1:] == 0)        # nothing written past the lines
    # an odd real line stride cannot be read as complex pairs
    xo = torch.zeros((rows, n + 1), dtype=torch.float32, device="cuda")
    with pytest.raises(ovx.OvxUnsupported):
        g(xo[:, :n], out=yp[:, : n // 2 + 1])


def test_c2r_ignores_imaginary_dc_and_nyquist(ovx):
    import torch
    rows, n = 4, 256
    rng = np.random.default_rng(6)
    xh = (rng.standard_normal((rows, n // 2 + 1)) + 1j * rng.standard_normal((rows, n // 2 + 1))).astype(np.complex64)
    clean = xh.copy()
    clean[:, 0] = clean[:, 0].real
    clean[:, n // 2] = clean[:, n // 2].real
    f = ovx.FftmReal(rows, n, ovx.ROW, ovx.FFT_INV)
    a = host(f(torch.from_numpy(xh).cuda()))
    b = host(f(torch.from_numpy(clean).cuda()))
    assert np.array_equal(a, b)
    assert o.rel_l2(a, n * np.fft.irfft(clean, n, axis=1)) < REL_L2


def test_real_fftm_refuses_aliased_input_and_output(ovx):
    """Input and output of a real Fftm differ in type and line length, so
    the transform runs out of place only; the same buffer on both sides is
    refused, not computed with overlapping reads and writes."""
    import ctypes
    import torch
    from openvsip_amd import _lib
    rows, n = 4, 64
    buf = torch.zeros(rows * (n + 2), dtype=torch.float32, device="cuda")
    f = ovx.FftmReal(rows, n, ovx.ROW, ovx.FFT_FWD, 1.0)
    p = ctypes.c_void_p(buf.data_ptr())
    st = _lib.lib().ovx_fftm_real_out_of_place(f._h, p, n + 2, 1, p, (n + 2) // 2, 1, rows, n,
                                               ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert st == _lib.OVX_ERR_UNSUPPORTED


def _gold_real():
    import os
    return np.load(os.path.join(os.path.dirname(__file__), "golden", "fft_ext_real.npz"))


@pytest.mark.parametrize("k", range(1, 9))
def test_r2c_reference_golden_vectors(ovx, k):
    """The reference's r2c golden files tests/fft_ext/data/rcfftop_f* (lengths
    4, 768, 1024; scales 1, 3, 1/1024), through a one-row real Fftm, at the
    reference's own criterion (tests/fft_ext/fft_ext.cpp:85-108: < -100 dB)."""
    import torch
    g = _gold_real()
    name = f"rcfftop_f{k}"
    x, e, sc = g[name + "_in"], g[name + "_expected"], float(g[name + "_scale"])
    f = ovx.FftmReal(1, len(x), ovx.ROW, ovx.FFT_FWD, sc)
    y = host(f(torch.from_numpy(x[None, :].copy()).cuda()))[0]
    say(name, len(x), sc, "db", o.diff_db(e.astype(np.complex128), y))
    assert o.diff_db(e.astype(np.complex128), y) < -100


@pytest.mark.parametrize("k", range(1, 8))
def test_c2r_reference_golden_vectors(ovx, k):
    """The reference's c2r golden files crfftop_f* (unnormalised inverse
    times the file's scale) through a one-row real inverse Fftm."""
    import torch
    g = _gold_real()
    name = f"crfftop_f{k}"
    xh, e, sc = g[name + "_in"], g[name + "_expected"], float(g[name + "_scale"])
    n = len(e)
    f = ovx.FftmReal(1, n, ovx.ROW, ovx.FFT_INV, sc)
    y = host(f(torch.from_numpy(xh[None, :].copy()).cuda()))[0]
    say(name, n, sc, "db", o.diff_db(e.astype(np.float64), y.astype(np.float64)))
    assert o.diff_db(e.astype(np.float64), y.astype(np.float64)) < -100
