"""Pins the CPU oracle (oracle/ovx_oracle.c) before GPU results are compared
with it.  Each check is independent of the formula it checks: golden vectors
from the reference's own tests, brute force, library routines (numpy.fft /
numpy.convolve), closed forms and invariants.  A planted mistake in an oracle
function fails at least one test here (see test_planted_mistakes_are_caught).
"""
import numpy as np
import pytest

import oracle as o

GOLD = None


def gold():
    global GOLD
    if GOLD is None:
        import os
        GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "fft_ext_cc.npz"))
    return GOLD


def cg(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def brute_dft(x, sign):
    """O(n^2) DFT by an explicit matrix, numpy float64 (independent of the C code)."""
    n = x.size
    k = np.arange(n)
    m = np.exp(sign * 2j * np.pi * np.outer(k, k) / n)
    return m @ x.astype(np.complex128)


# ---- golden vectors of the reference (tests/fft_ext/data/ccfftop_f1..f8) ----

@pytest.mark.parametrize("k", range(1, 9))
def test_dft_ref_matches_reference_golden(k):
    g = gold()
    name = f"ccfftop_f{k}"
    x, e, s = g[name + "_in"], g[name + "_expected"], float(g[name + "_scale"])
    y = o.dft_ref(x, o.FWD, s)
    # the reference's own acceptance threshold is < -100 dB (fft_ext.cpp:95)
    assert o.diff_db(y, e.astype(np.complex128)) < -100


@pytest.mark.parametrize("k", range(1, 9))
def test_fft64_matches_reference_golden(k):
    g = gold()
    name = f"ccfftop_f{k}"
    x, e, s = g[name + "_in"], g[name + "_expected"], float(g[name + "_scale"])
    y = o.fft64(x.astype(np.complex128)) * s
    assert o.diff_db(e, y) < -100


# ---- real<->complex golden vectors (rcfftop_f1..f8, crfftop_f1..f7) ---------

def gold_real():
    import os
    return np.load(os.path.join(os.path.dirname(__file__), "golden", "fft_ext_real.npz"))


@pytest.mark.parametrize("k", range(1, 9))
def test_rfftm64_matches_reference_golden(k):
    """The real-input transform of the oracle against the reference's own
    r2c golden files (tests/fft_ext/fft_ext.cpp:213-262, < -100 dB)."""
    g = gold_real()
    name = f"rcfftop_f{k}"
    x, e, s = g[name + "_in"], g[name + "_expected"], float(g[name + "_scale"])
    y = o.rfftm64(x[None, :], s)[0]
    assert o.diff_db(e.astype(np.complex128), y) < -100


@pytest.mark.parametrize("k", range(1, 8))
def test_irfftm64_matches_reference_golden(k):
    """The oracle's complex-to-real transform (unnormalised, times scale)
    against the reference's c2r golden files (fft_ext.cpp:164-210)."""
    g = gold_real()
    name = f"crfftop_f{k}"
    xh, e, s = g[name + "_in"], g[name + "_expected"], float(g[name + "_scale"])
    y = o.irfftm64(xh[None, :], len(e), s)[0]
    assert o.diff_db(e.astype(np.float64), y) < -100


# ---- FP64 FFT against brute force and a library routine ----------------------

@pytest.mark.parametrize("n", list(range(1, 65)) + [96, 100, 243, 250, 252, 768, 1000])
@pytest.mark.parametrize("sign", [o.FWD, o.INV])
def test_fft64_matches_brute_force(n, sign):
    rng = np.random.default_rng(n)
    x = cg(rng, n).astype(np.complex128)
    y = o.fft64(x, sign)
    ref = brute_dft(x, sign)
    assert np.abs(y - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("n", [4096, 12288, 16384, 65536, 262144])
def test_fft64_matches_numpy_large(n):
    rng = np.random.default_rng(7)
    x = cg(rng, n).astype(np.complex128)
    assert o.rel_l2(o.fft64(x, o.FWD), np.fft.fft(x)) < 1e-14
    assert o.rel_l2(o.fft64(x, o.INV), np.fft.ifft(x) * n) < 1e-14


def test_fft64_closed_forms():
    n = 1024
    imp = np.zeros(n, np.complex128)
    imp[0] = 1
    assert np.allclose(o.fft64(imp), np.ones(n), atol=1e-15)
    assert np.allclose(o.fft64(np.ones(n, np.complex128)), n * imp, atol=1e-9)
    k0 = 37
    tone = np.exp(2j * np.pi * k0 * np.arange(n) / n)
    y = o.fft64(tone)
    expect = np.zeros(n, np.complex128)
    expect[k0] = n
    assert np.abs(y - expect).max() < 1e-9
    rng = np.random.default_rng(1)
    x = cg(rng, n).astype(np.complex128)
    y = o.fft64(x)
    assert abs(np.vdot(y, y).real - n * np.vdot(x, x).real) < 1e-9 * n * np.vdot(x, x).real  # Parseval
    assert o.rel_l2(o.fft64(y, o.INV) / n, x) < 1e-15


def test_fftm64_rows_and_scale():
    rng = np.random.default_rng(2)
    x = cg(rng, (5, 768)).astype(np.complex128)
    y = o.fftm64(x, o.INV, 1.0 / 768)
    assert o.rel_l2(y, np.fft.ifft(x, axis=1)) < 1e-14


def test_dftm_ref_rows():
    rng = np.random.default_rng(3)
    x = cg(rng, (3, 64))
    y = o.dftm_ref(x, o.FWD, 2.0)
    for r in range(3):
        assert o.rel_l2(y[r], 2.0 * brute_dft(x[r], -1)) < 1e-6


# ---- fast convolution -------------------------------------------------------

@pytest.mark.parametrize("n,shift", [(16, 2), (64, 2), (2048, 5)])
def test_fastconv64_shift_kernel(n, shift):
    """tests/fastconv.cpp:307-330: weights = impulse at `shift` times 2 ->
    output row = 2 * input row shifted by `shift` (< -100 dB)."""
    scale = 2.0
    weights = np.zeros(n, np.complex64)
    weights[shift] = scale
    w = np.fft.fft(weights.astype(np.complex128)).astype(np.complex64)  # t2f
    x = np.tile(np.arange(n, dtype=np.float32), (16, 1)).astype(np.complex64)  # ramp(0, 1, n)
    y = o.fastconv64(x, w)
    assert o.diff_db(scale * x[:, : n - shift], y[:, shift:]) < -100


def test_fastconv64_matches_brute_circular_convolution():
    rng = np.random.default_rng(4)
    n = 48
    x = cg(rng, (3, n))
    h = cg(rng, n)
    w = np.fft.fft(h.astype(np.complex128)).astype(np.complex64)
    y = o.fastconv64(x, w)
    # circular convolution with the time-domain kernel whose spectrum is w
    hh = np.fft.ifft(w.astype(np.complex128))
    ref = np.array([[sum(x[r, (i - j) % n] * hh[j] for j in range(n)) for i in range(n)] for r in range(3)])
    assert o.rel_l2(y, ref) < 1e-12


def test_fastconv64_matrix_coefficients():
    """tests/fastconv.cpp:215-245 (fused_m_multi): a coefficient per row."""
    rng = np.random.default_rng(5)
    x = cg(rng, (4, 256))
    c = cg(rng, (4, 256))
    y = o.fastconv64(x, c)
    ref = np.fft.ifft(np.fft.fft(x.astype(np.complex128), axis=1) * c.astype(np.complex128), axis=1)
    assert o.rel_l2(y, ref) < 1e-14


def test_fastconv_ref_dft_and_port_agree_with_fp64():
    rng = np.random.default_rng(6)
    x = cg(rng, (4, 512))
    w = cg(rng, 512)
    y64 = o.fastconv64(x, w)
    assert o.rel_l2(o.fastconv_ref_dft(x, w), y64) < 2e-6
    assert o.rel_l2(o.fastconv_port(x, w), y64) < 2e-6
    assert o.rel_l2(o.fastconv_port(x, w, threads=3), y64) < 2e-6


@pytest.mark.parametrize("n", [768, 3072, 4096])
def test_fftm_port_accuracy(n):
    rng = np.random.default_rng(n)
    x = cg(rng, (2, n))
    y = o.fftm_port(x, o.FWD)
    assert o.rel_l2(y, np.fft.fft(x.astype(np.complex128), axis=1)) < 1e-6


# ---- correlation ------------------------------------------------------------

def corr_independent(bias, support, ref, x):
    """Correlation from one linear convolution (a derivation independent of
    the position loop): out[n] = (conj(x) * reverse(ref))[n + M-1-shift]; the
    unbiased divisor is the number of overlapping terms, except the C-VSIPL
    same-support end rule for even M (src/ovxx/signal/corr.hpp:40-65)."""
    m, n = ref.size, x.size
    shift = {"full": m - 1, "same": m // 2, "min": 0}[support]
    p = {"full": n + m - 1, "same": n, "min": n - m + 1}[support]
    c = np.convolve(np.conj(x.astype(np.complex128)), ref[::-1].astype(np.complex128))
    off = m - 1 - shift
    out = c[off: off + p].copy()
    if bias == "unbiased":
        for i in range(p):
            cnt = sum(1 for k in range(m) if 0 <= i + k - shift < n)
            if support == "same" and m % 2 == 0 and i >= n - m // 2:
                cnt -= 1
            out[i] /= cnt
    return out


@pytest.mark.parametrize("support", ["full", "same", "min"])
@pytest.mark.parametrize("bias", ["biased", "unbiased"])
@pytest.mark.parametrize("m,n", [(1, 1), (4, 32), (7, 32), (16, 16), (17, 40)])
def test_corr_oracles_match_independent_derivation(support, bias, m, n):
    rng = np.random.default_rng(m * 100 + n)
    ref, x = cg(rng, m), cg(rng, n)
    ind = corr_independent(bias, support, ref, x)
    assert o.rel_l2(o.corr64(bias, support, ref, x), ind) < 1e-12
    assert o.rel_l2(o.corr_ref(bias, support, ref, x), ind) < 1e-5


# ---- convolution ------------------------------------------------------------

@pytest.mark.parametrize("support", ["full", "same", "min"])
@pytest.mark.parametrize("d", [1, 2, 3])
@pytest.mark.parametrize("m,n", [(1, 8), (4, 32), (5, 33), (9, 9)])
def test_conv_oracles_match_numpy(support, d, m, n):
    rng = np.random.default_rng(m * 7 + n + d)
    k, x = cg(rng, m), cg(rng, n)
    c = np.convolve(x.astype(np.complex128), k.astype(np.complex128))
    off = {"full": 0, "same": m // 2, "min": m - 1}[support]
    p = o.conv_output_size(support, m, n, d)
    ind = np.array([c[i * d + off] for i in range(p)])
    assert o.rel_l2(o.conv64(support, k, x, d), ind) < 1e-12
    assert o.rel_l2(o.conv_ref(support, k, x, d), ind) < 1e-5


def test_conv_output_sizes_closed_form():
    # conv.hpp:103-121 with the "correct min size" rule: min size = ceil((N-M+1)/D)
    for m in range(1, 6):
        for n in range(m, 20):
            for d in range(1, 4):
                assert o.conv_output_size("full", m, n, d) == len(range(0, n + m - 1, d))
                assert o.conv_output_size("same", m, n, d) == len(range(0, n, d))
                assert o.conv_output_size("min", m, n, d) == len(range(0, n - m + 1, d))


def test_conv_kernel_symmetry():
    c = np.array([1, 2, 3], np.complex64)
    assert list(o.conv_kernel("nonsym", c)) == [1, 2, 3]
    assert list(o.conv_kernel("sym_even_len_odd", c)) == [1, 2, 3, 2, 1]
    assert list(o.conv_kernel("sym_even_len_even", c)) == [1, 2, 3, 3, 2, 1]


# ---- metric -----------------------------------------------------------------

def test_diff_db_hand_values():
    a = np.array([1 + 0j, 0j], np.complex64)
    b = np.array([1 + 0j, 0.001 + 0j])
    # |err|^2 = 1e-6, refmax = 1 -> 10 log10(1e-6 / 2)
    assert abs(o.diff_db(a, b) - 10 * np.log10(0.5e-6)) < 1e-6
    assert o.diff_db(a, a.astype(np.complex128)) == -201.0
    assert o.diff_db(np.array([np.nan], np.complex64), np.array([0j])) == 201.0


# ---- the checks above catch plausible mistakes ------------------------------

def test_planted_mistakes_are_caught():
    """Mistakes planted in a copy of the oracle's results (on the CPU) fail the
    same comparisons the tests use: a dropped conjugate, an off-by-one shift,
    a wrong sign and a missing scale."""
    rng = np.random.default_rng(11)
    ref, x = cg(rng, 7), cg(rng, 32)
    good = o.corr64("biased", "full", ref, x)
    ind = corr_independent("biased", "full", ref, x)
    assert o.rel_l2(good, ind) < 1e-12
    no_conj = np.convolve(x.astype(np.complex128), ref[::-1].astype(np.complex128))
    assert o.rel_l2(no_conj, ind) > 1e-2
    assert o.rel_l2(np.roll(good, 1), ind) > 1e-2
    n = 64
    xx = cg(rng, n).astype(np.complex128)
    assert o.rel_l2(o.fft64(xx, o.INV), np.fft.fft(xx)) > 1e-2       # wrong sign
    w = cg(rng, n)
    y = o.fastconv64(xx[None, :].astype(np.complex64), w, scale=1.0)  # missing 1/N
    ref_fc = np.fft.ifft(np.fft.fft(xx) * w)
    assert o.rel_l2(y[0], ref_fc) > 1e-2


# ---- real <-> complex Fftm ----------------------------------------------------

@pytest.mark.parametrize("n", [2, 4, 6, 16, 30, 64, 1000, 4096])
def test_real_fftm_oracle_matches_numpy(n):
    """r2c against numpy.fft.rfft; c2r against n * numpy.fft.irfft (numpy's
    inverse is normalised, FFTW's c2r is not; both drop the imaginary parts of
    frequencies 0 and n/2)."""
    rng = np.random.default_rng(n)
    x = rng.standard_normal((3, n))
    xh = o.rfftm64(x)
    assert xh.shape == (3, n // 2 + 1)
    assert o.rel_l2(xh, np.fft.rfft(x, axis=1)) < 1e-13
    y = rng.standard_normal((3, n // 2 + 1)) + 1j * rng.standard_normal((3, n // 2 + 1))
    assert o.rel_l2(o.irfftm64(y, n), n * np.fft.irfft(y, n, axis=1)) < 1e-13
    # round trip: c2r(r2c(x)) = n x
    assert o.rel_l2(o.irfftm64(xh, n), n * x) < 1e-13


def test_real_fftm_oracle_closed_forms():
    n = 32
    imp = np.zeros((1, n))
    imp[0, 0] = 1.0
    assert np.allclose(o.rfftm64(imp), np.ones((1, n // 2 + 1)))      # impulse -> flat
    const = np.ones((1, n))
    ref = np.zeros((1, n // 2 + 1), complex)
    ref[0, 0] = n
    assert np.allclose(o.rfftm64(const), ref)                         # constant -> DC
    k0 = 5
    tone = np.cos(2 * np.pi * k0 * np.arange(n) / n)[None, :]
    ref = np.zeros((1, n // 2 + 1), complex)
    ref[0, k0] = n / 2
    assert np.allclose(o.rfftm64(tone), ref, atol=1e-12)             # cosine -> one bin
    # a planted mistake (Hermitian extension without the conjugate) fails
    y = np.random.default_rng(2).standard_normal((1, n // 2 + 1)) * (1 + 1j)
    bad = o.hermitian_extend(y, n)
    bad[:, n // 2 + 1:] = np.conj(bad[:, n // 2 + 1:])
    assert o.rel_l2(o.fftm64(bad, o.INV).real, n * np.fft.irfft(y, n, axis=1)) > 1e-2
