"""CPU oracle for OpenVSIP's fast-convolution path -- TEST INFRASTRUCTURE ONLY.

This package restates the reference's CPU algorithms (see ovx_oracle.c for the
file:line each function follows).  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import it, and only as the values GPU results
are compared with or as the timed CPU baseline.  The product (openvsip_amd, libovxhip.so) never uses it.

Pinning status of each function is listed in the header of ovx_oracle.c; every
function has at least one independent check in tests/test_oracle.py.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

FWD = -1
INV = +1

_f32p = np.ctypeslib.ndpointer(dtype=np.complex64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")
_L = ctypes.c_long
_D = ctypes.c_double
_I = ctypes.c_int


def build():
    """Compile liboracle.so with make (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        l = ctypes.CDLL(_LIB_PATH)
        l.ovo_dft_ref.argtypes = [_f32p, _L, _f32p, _L, _L, _I, _D]
        l.ovo_dftm_ref.argtypes = [_f32p, _L, _f32p, _L, _L, _L, _I, _D]
        l.ovo_fft_f64.argtypes = [_f64p, _f64p, _L, _I]
        l.ovo_fftm_f64.argtypes = [_f64p, _L, _f64p, _L, _L, _L, _I, _D]
        l.ovo_fastconv_f64.argtypes = [_f32p, _L, _f32p, _L, _f64p, _L, _L, _D]
        l.ovo_fastconv_ref_dft.argtypes = [_f32p, _L, _f32p, _L, _f32p, _L, _L, _L, _D]
        l.ovo_fastconv_port_f32.argtypes = [_f32p, _L, _f32p, _L, _f32p, _L, _L, _L, _D, _I]
        l.ovo_fastconv_port_f32.restype = _I
        l.ovo_fftm_port_f32.argtypes = [_f32p, _L, _f32p, _L, _L, _L, _I, _D]
        l.ovo_corr_output_size.argtypes = [_I, _L, _L]
        l.ovo_corr_output_size.restype = _L
        l.ovo_corr_ref_f32.argtypes = [_I, _I, _f32p, _L, _f32p, _L, _f32p]
        l.ovo_corr_f64.argtypes = [_I, _I, _f32p, _L, _f32p, _L, _f64p]
        l.ovo_conv_output_size.argtypes = [_I, _L, _L, _L]
        l.ovo_conv_output_size.restype = _L
        l.ovo_conv_kernel.argtypes = [_I, _f32p, _L, _f32p]
        l.ovo_conv_kernel.restype = _L
        l.ovo_conv_ref_f32.argtypes = [_I, _f32p, _L, _f32p, _L, _L, _f32p]
        l.ovo_conv_f64.argtypes = [_I, _f32p, _L, _f32p, _L, _L, _f64p]
        l.ovo_diff_db.argtypes = [_f32p, _f64p, _L]
        l.ovo_diff_db.restype = _D
        _lib = l
    return _lib


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.complex64)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


# ---- transforms -----------------------------------------------------------

def dft_ref(x, direction=FWD, scale=1.0):
    """Reference builtin-DFT transform of one vector (float in/out)."""
    x = _c64(x)
    out = np.empty_like(x)
    lib().ovo_dft_ref(x, 1, out, 1, x.size, direction, scale)
    return out


def dftm_ref(x, direction=FWD, scale=1.0):
    x = _c64(x)
    out = np.empty_like(x)
    r, n = x.shape
    lib().ovo_dftm_ref(x, n, out, n, r, n, direction, scale)
    return out


def fft64(x, direction=FWD):
    x = _c128(x)
    out = np.empty_like(x)
    lib().ovo_fft_f64(x, out, x.size, direction)
    return out


def fftm64(x, direction=FWD, scale=1.0):
    x = _c128(np.atleast_2d(x))
    out = np.empty_like(x)
    r, n = x.shape
    lib().ovo_fftm_f64(x, n, out, n, r, n, direction, scale)
    return out


def rfftm64(x, scale=1.0):
    """Real-input Fftm by rows (FFTW r2c as the reference's real fftm calls it,
    src/ovxx/fftw/fftw.cpp:778-843): n real samples -> n/2 + 1 complex
    values, FP64, times scale."""
    x = np.atleast_2d(np.asarray(x, dtype=np.float64))
    n = x.shape[1]
    return fftm64(x.astype(np.complex128))[:, : n // 2 + 1] * scale


def hermitian_extend(xh, n):
    """Full n-point spectrum from the half spectrum X[0 .. n/2] of a real
    signal (imaginary parts of frequencies 0 and n/2 dropped, as FFTW's c2r
    and the reference's real inverse Fftm assume them zero)."""
    xh = np.atleast_2d(np.asarray(xh, dtype=np.complex128)).copy()
    h = n // 2
    xh[:, 0] = xh[:, 0].real
    xh[:, h] = xh[:, h].real
    full = np.empty((xh.shape[0], n), dtype=np.complex128)
    full[:, : h + 1] = xh[:, : h + 1]
    full[:, h + 1:] = np.conj(xh[:, 1:h][:, ::-1])
    return full


def irfftm64(xh, n, scale=1.0):
    """Complex-to-real Fftm by rows (FFTW c2r, src/ovxx/fftw/fftw.cpp:845-903),
    unnormalised, FP64, times scale."""
    return fftm64(hermitian_extend(xh, n), INV).real * scale


def fftm_port(x, direction=FWD, scale=1.0):
    x = _c64(np.atleast_2d(x))
    out = np.empty_like(x)
    r, n = x.shape
    lib().ovo_fftm_port_f32(x, n, out, n, r, n, direction, scale)
    return out


# ---- fast convolution -------------------------------------------------------

def fastconv64(x, w, scale=None):
    """FP64 fastconv of float inputs; w is a vector (vmmul<0>) or a matrix."""
    x = _c64(np.atleast_2d(x))
    w = _c64(w)
    r, n = x.shape
    ldw = 0 if w.ndim == 1 else n
    if scale is None:
        scale = 1.0 / n
    y = np.empty((r, n), dtype=np.complex128)
    lib().ovo_fastconv_f64(x, n, w, ldw, y, r, n, scale)
    return y


def fastconv_ref_dft(x, w, scale=None):
    """The reference's fastconv on its builtin O(N^2) dft backend (float out)."""
    x = _c64(np.atleast_2d(x))
    w = _c64(w)
    r, n = x.shape
    ldw = 0 if w.ndim == 1 else n
    if scale is None:
        scale = 1.0 / n
    y = np.empty((r, n), dtype=np.complex64)
    lib().ovo_fastconv_ref_dft(x, n, w, ldw, y, n, r, n, scale)
    return y


def fastconv_port(x, w, scale=None, threads=1):
    """FP32 four-pass CPU fastconv (the CPU baseline)."""
    x = _c64(np.atleast_2d(x))
    w = _c64(w)
    r, n = x.shape
    ldw = 0 if w.ndim == 1 else n
    if scale is None:
        scale = 1.0 / n
    y = np.empty((r, n), dtype=np.complex64)
    lib().ovo_fastconv_port_f32(x, n, w, ldw, y, n, r, n, scale, threads)
    return y


# ---- correlation / convolution ---------------------------------------------

SUPPORT = {"full": 0, "same": 1, "min": 2}
BIAS = {"biased": 0, "unbiased": 1}
SYMMETRY = {"nonsym": 0, "sym_even_len_odd": 1, "sym_even_len_even": 2}


def corr_output_size(support, m, n):
    return lib().ovo_corr_output_size(SUPPORT[support], m, n)


def corr_ref(bias, support, ref, x):
    ref, x = _c64(ref), _c64(x)
    p = corr_output_size(support, ref.size, x.size)
    out = np.empty(p, dtype=np.complex64)
    lib().ovo_corr_ref_f32(BIAS[bias], SUPPORT[support], ref, ref.size, x, x.size, out)
    return out


def corr64(bias, support, ref, x):
    ref, x = _c64(ref), _c64(x)
    p = corr_output_size(support, ref.size, x.size)
    out = np.empty(p, dtype=np.complex128)
    lib().ovo_corr_f64(BIAS[bias], SUPPORT[support], ref, ref.size, x, x.size, out)
    return out


def conv_output_size(support, m, n, decimation=1):
    return lib().ovo_conv_output_size(SUPPORT[support], m, n, decimation)


def conv_kernel(symmetry, coeff):
    coeff = _c64(coeff)
    k = np.empty(2 * coeff.size, dtype=np.complex64)
    m = lib().ovo_conv_kernel(SYMMETRY[symmetry], coeff, coeff.size, k)
    return k[:m].copy()


def conv_ref(support, kernel, x, decimation=1):
    kernel, x = _c64(kernel), _c64(x)
    p = conv_output_size(support, kernel.size, x.size, decimation)
    out = np.empty(p, dtype=np.complex64)
    lib().ovo_conv_ref_f32(SUPPORT[support], kernel, kernel.size, x, x.size, decimation, out)
    return out


def conv64(support, kernel, x, decimation=1):
    kernel, x = _c64(kernel), _c64(x)
    p = conv_output_size(support, kernel.size, x.size, decimation)
    out = np.empty(p, dtype=np.complex128)
    lib().ovo_conv_f64(SUPPORT[support], kernel, kernel.size, x, x.size, decimation, out)
    return out


# ---- metrics ----------------------------------------------------------------

def diff_db(a, b):
    """The reference's test::diff in dB (support/test/diff.hpp:62-96)."""
    a = _c64(np.ravel(a))
    b = _c128(np.ravel(b))
    return lib().ovo_diff_db(a, b, a.size)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.complex128).ravel()
    b = np.asarray(b, dtype=np.complex128).ravel()
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))
