"""GPU parity tests: the HIP kernels (through the C-ABI) against the CPU
oracle on the same seeded inputs.  Tolerance for complex<float>: rel-L2 <= 1e-5
against the FP64 oracle (the tolerance set for complex<float> results of
the reference's CPU fastconv), plus the reference's own test::diff < -100 dB criterion
(support/test/diff.hpp:62-96, tests/fastconv.cpp:327)."""
import numpy as np
import pytest

import oracle as o
from openvsip_amd import workloads as wl

pytestmark = pytest.mark.gpu

REL_L2 = 1e-5
DB = -100.0


def cg(rng, shape):
    return wl.cgauss(rng, shape)


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def say(*a):
    print(*a, flush=True)


POW2 = [2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384]
SMOOTH = [3, 6, 12, 24, 48, 96, 768, 1536, 3072, 6144, 12288, 5, 10, 20, 40, 80, 640, 5120, 10240,
          9, 18, 36, 72, 144, 576, 2304, 9216]


@pytest.mark.parametrize("n", POW2 + SMOOTH)
@pytest.mark.parametrize("direction", ["fwd", "inv"])
def test_fftm_rows_vs_fp64(ovx, n, direction):
    say("fftm", n, direction)
    rows = 37  # ragged: not a multiple of the rows per block
    rng = np.random.default_rng(n)
    x = cg(rng, (rows, n))
    d = ovx.FFT_FWD if direction == "fwd" else ovx.FFT_INV
    scale = 1.0 if direction == "fwd" else 1.0 / n
    f = ovx.Fftm(rows, n, ovx.ROW, d, scale)
    y = host(f(dev(x)))
    ref = o.fftm64(x.astype(np.complex128), o.FWD if direction == "fwd" else o.INV, scale)
    err = o.rel_l2(y, ref)
    say("  rel_l2", err, "db", o.diff_db(y, ref))
    assert err < REL_L2
    assert o.diff_db(y, ref) < DB


def test_fftm_strided_rows_and_in_place(ovx):
    import torch
    n, rows, ld = 1024, 20, 1100
    rng = np.random.default_rng(9)
    big = cg(rng, (rows, ld))
    t = dev(big)
    f = ovx.Fftm(rows, n, ovx.ROW, ovx.FFT_FWD, 1.0)
    view = t[:, :n]
    out = torch.zeros((rows, ld), dtype=torch.complex64, device="cuda")
    f(view, out[:, :n])
    y = host(out)
    ref = o.fftm64(big[:, :n].astype(np.complex128), o.FWD)
    assert o.rel_l2(y[:, :n], ref) < REL_L2
    assert np.all(y[:, n:] == 0)  # nothing written past the row
    f.in_place(view)
    y2 = host(t)
    assert o.rel_l2(y2[:, :n], ref) < REL_L2
    assert np.array_equal(y2[:, n:], big[:, n:])


def test_fftm_column_axis(ovx):
    """Fftm along columns (A = col) on column-major data (the layout
    fftm_backend::query_layout asks for, backend.hpp:713-723)."""
    import torch
    rows, cols = 512, 24
    rng = np.random.default_rng(10)
    x = cg(rng, (rows, cols))
    # column-major storage: element (r, c) at c*rows + r
    xt = dev(np.ascontiguousarray(x.T))
    xcm = xt.t()  # shape (rows, cols), strides (1, rows)
    f = ovx.Fftm(rows, cols, ovx.COL, ovx.FFT_FWD, 1.0)
    out_t = torch.empty((cols, rows), dtype=torch.complex64, device="cuda")
    f(xcm, out_t.t())
    y = host(out_t).T
    ref = np.fft.fft(x.astype(np.complex128), axis=0)
    assert o.rel_l2(y, ref) < REL_L2


@pytest.mark.parametrize("k", [1, 2, 3, 5, 6, 7, 8])
def test_fft_reference_golden_vectors(ovx, k):
    """The reference's own golden files tests/fft_ext/data/ccfftop_f*
    (committed as data in tests/golden/fft_ext_cc.npz), < -100 dB as
    tests/fft_ext/fft_ext.cpp:95 requires."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "fft_ext_cc.npz"))
    name = f"ccfftop_f{k}"
    x, e, s = g[name + "_in"], g[name + "_expected"], float(g[name + "_scale"])
    say("golden", name, x.size, s)
    f = ovx.Fft(x.size, ovx.FFT_FWD, s)
    y = host(f(dev(x)))
    assert o.diff_db(y, e.astype(np.complex128)) < DB


def test_unsupported_length_is_reported(ovx):
    """A length the backend does not handle (19 = ccfftop_f4) is refused with
    OVX_ERR_UNSUPPORTED so the dispatcher can fall through to the next backend
    (the reference's rt_valid() == false)."""
    from openvsip_amd import OvxUnsupported
    with pytest.raises(OvxUnsupported):
        ovx.Fft(19, ovx.FFT_FWD, 1.0)


# ---- fused fast convolution -------------------------------------------------

@pytest.mark.parametrize("n,shift,rows", [(16, 2, 16), (64, 2, 16), (64, 2, 64), (2048, 7, 16)])
def test_fastconv_shift_kernel(ovx, n, shift, rows):
    """tests/fastconv.cpp:307-330 / 341-365: a shift kernel (impulse of 2 at
    `shift`) turns each ramp row into 2 x the row shifted by `shift`."""
    say("shift", n, shift, rows)
    weights = np.zeros(n, np.complex64)
    weights[shift] = 2.0
    w = np.fft.fft(weights.astype(np.complex128)).astype(np.complex64)
    x = np.tile(np.arange(n, dtype=np.float32), (rows, 1)).astype(np.complex64)
    fc = ovx.Fastconv(rows, n)
    y = host(fc(dev(x), dev(w)))
    assert o.diff_db(2.0 * x[:, : n - shift], y[:, shift:].astype(np.complex128)) < DB


def test_fastconv_pulse_compression_small(ovx):
    """64 x 2048, seed 0, chirp B*T = 512 (the reference CPU-runnable case)."""
    x, w = wl.pulse_compression_small()
    fc = ovx.Fastconv(64, 2048)
    y = host(fc(dev(x), dev(w)))
    ref = o.fastconv64(x, w)
    err = o.rel_l2(y, ref)
    say("64x2048 rel_l2", err, "max_abs", np.abs(y - ref).max(), "db", o.diff_db(y, ref))
    assert err < REL_L2
    assert o.diff_db(y, ref) < DB


def test_fastconv_pulse_compression_small_vs_reference_fixture(ovx):
    """Against rows computed by the reference's own dft-backend fastconv
    semantics (oracle.fastconv_ref_dft: float-rounded forward dftm, float
    vmmul, float-rounded inverse dftm, float scale pass)."""
    x, w = wl.pulse_compression_small()
    rows = [0, 31, 63]
    ref = o.fastconv_ref_dft(x[rows], w)
    fc = ovx.Fastconv(64, 2048)
    y = host(fc(dev(x), dev(w)))[rows]
    assert o.rel_l2(y, ref) < REL_L2


def test_fastconv_radar_256x16384(ovx):
    x, w, delays, amps = wl.radar(256, 16384)
    fc = ovx.Fastconv(256, 16384)
    y = host(fc(dev(x), dev(w)))
    ref = o.fastconv64(x, w)
    err = o.rel_l2(y, ref)
    say("256x16384 rel_l2", err, "max_abs", np.abs(y - ref).max())
    assert err < REL_L2
    assert o.diff_db(y, ref) < DB


def test_fastconv_matched_filter_peaks(ovx):
    """LFM matched filter (W = conj(FFT(chirp))): the compressed peak of each
    echo lands at its delay with amplitude ~ echo amplitude x chirp energy."""
    x, _, delays, amps = wl.radar(8, 16384, seed=5)
    chirp = wl.lfm_chirp(4096, 1024)
    wm = np.conj(wl.replica_spectrum(chirp, 16384)).astype(np.complex64)
    fc = ovx.Fastconv(8, 16384)
    y = host(fc(dev(x), dev(wm)))
    energy = 4096.0
    for r in range(8):
        for d, a in zip(delays[r], amps[r]):
            peak = np.abs(y[r, d])
            assert abs(peak - a * energy) < 0.2 * energy, (r, d, a, peak)


# Error bound of the BF16x3 variant against FP64.  Each operand is split as
# hi + lo with |x - hi - lo| <= 2^-16 |x| (bf16 keeps 8 significant bits) and
# the dropped lo.lo product is of the same order, so every radix-16 stage adds
# about 1e-5 relative error (RMS); six such stages plus the FP32 steps stay
# below this bound.  It is the precision variant's stated bound, looser than
# the FP32 path's parity tolerance.
BF16X3_REL_L2 = 1e-4


def test_fastconv_bf16x3_mfma_vs_fp64(ovx):
    x, w, delays, amps = wl.radar(64, 16384, seed=7)
    ref = o.fastconv64(x, w)
    fc32 = ovx.Fastconv(64, 16384)
    fcb = ovx.Fastconv(64, 16384, precision="bf16x3")
    assert fcb.name == "fastconv:bf16x3-mfma"
    xd, wd = dev(x), dev(w)
    y32 = host(fc32(xd, wd))
    yb = host(fcb(xd, wd))
    e32, eb = o.rel_l2(y32, ref), o.rel_l2(yb, ref)
    say("bf16x3 rel_l2", eb, "max_abs", np.abs(yb - ref).max(), "| fp32 rel_l2", e32,
        "max_abs", np.abs(y32 - ref).max())
    assert e32 < REL_L2
    assert eb < BF16X3_REL_L2
    # ragged rows, strided input and in-place output
    xs = np.zeros((5, 16384 + 64), np.complex64)
    xs[:, :16384] = x[:5]
    xsd = dev(xs)
    fcb(xsd[:, :16384], wd, out=xsd[:, :16384], rows=5)
    assert o.rel_l2(host(xsd)[:, :16384], ref[:5]) < BF16X3_REL_L2
    # a spectrum matrix is refused
    with pytest.raises(ovx.OvxUnsupported):
        fcb(xd, dev(np.repeat(w[None, :], 64, axis=0)))


def test_fastconv_coefficient_matrix_and_in_place(ovx):
    rows, n = 33, 4096
    rng = np.random.default_rng(12)
    x = cg(rng, (rows, n))
    c = cg(rng, (rows, n))
    fc = ovx.Fastconv(rows, n)
    t = dev(x)
    fc(t, dev(c), out=t)  # in place
    y = host(t)
    ref = o.fastconv64(x, c)
    assert o.rel_l2(y, ref) < REL_L2


@pytest.mark.parametrize("n", [512, 1024, 3072, 6144, 8192, 12288, 10240, 18, 4608])
def test_fastconv_sizes(ovx, n):
    rows = 19
    rng = np.random.default_rng(n + 1)
    x = cg(rng, (rows, n))
    w = wl.replica_spectrum(wl.lfm_chirp(n, n // 8), n)
    fc = ovx.Fastconv(rows, n)
    y = host(fc(dev(x), dev(w)))
    err = o.rel_l2(y, o.fastconv64(x, w))
    say("fastconv", n, err)
    assert err < REL_L2


def test_replica_resident_spectrum(ovx):
    n, m = 16384, 4096
    chirp = wl.lfm_chirp(m, 1024).astype(np.complex64)
    rep = ovx.Replica(n, chirp)
    spec = host(rep.spectrum())
    pad = np.zeros(n, np.complex128)
    pad[:m] = chirp
    assert o.rel_l2(spec, np.fft.fft(pad)) < REL_L2


# ---- correlation / convolution ------------------------------------------------

@pytest.mark.parametrize("support", ["full", "same", "min"])
@pytest.mark.parametrize("bias", ["biased", "unbiased"])
@pytest.mark.parametrize("m,n", [(1, 5), (4, 32), (7, 100), (16, 16), (100, 1000), (1024, 65536)])
def test_correlation_vs_oracle(ovx, support, bias, m, n):
    say("corr", support, bias, m, n)
    rng = np.random.default_rng(m + n)
    ref, x = cg(rng, m), cg(rng, n)
    c = ovx.Correlation(support, m, n)
    y = host(c(bias, dev(ref), dev(x)))
    exp = o.corr64(bias, support, ref, x)
    err = o.rel_l2(y, exp)
    assert err < REL_L2, err
    assert o.diff_db(y, exp) < DB


@pytest.mark.parametrize("symmetry", ["nonsym", "sym_even_len_odd", "sym_even_len_even"])
@pytest.mark.parametrize("support", ["full", "same", "min"])
@pytest.mark.parametrize("d", [1, 3])
def test_convolution_vs_oracle(ovx, symmetry, support, d):
    rng = np.random.default_rng(d + len(support) + len(symmetry))
    coeff, x = cg(rng, 9), cg(rng, 3000)
    conv = ovx.Convolution(coeff, symmetry, support, 3000, d)
    y = host(conv(dev(x)))
    k = o.conv_kernel(symmetry, coeff)
    exp = o.conv64(support, k, x, d)
    assert y.shape == exp.shape
    assert o.rel_l2(y, exp) < REL_L2


# ---- rows longer than LDS: two-pass four-step ----------------------------------

@pytest.mark.parametrize("n", [32768, 65536, 131072, 262144])
@pytest.mark.parametrize("direction", ["fwd", "inv"])
def test_fftm_fourstep_vs_fp64(ovx, n, direction):
    say("fftm four-step", n, direction)
    rows = 3
    rng = np.random.default_rng(n + 3)
    x = cg(rng, (rows, n))
    d = ovx.FFT_FWD if direction == "fwd" else ovx.FFT_INV
    scale = 1.0 if direction == "fwd" else 1.0 / n
    f = ovx.Fftm(rows, n, ovx.ROW, d, scale)
    assert "four-step" in f.name
    y = host(f(dev(x)))
    ref = o.fftm64(x.astype(np.complex128), o.FWD if direction == "fwd" else o.INV, scale)
    err = o.rel_l2(y, ref)
    say("  rel_l2", err)
    assert err < REL_L2


@pytest.mark.parametrize("n", [65536, 262144])
def test_fastconv_fourstep_vs_fp64(ovx, n):
    rows = 64 if n == 262144 else 20
    say("fastconv four-step", rows, n)
    x, w, _, _ = wl.radar(rows, n, seed=1)
    fc = ovx.Fastconv(rows, n)
    assert "four-step" in fc.name
    y = host(fc(dev(x), dev(w)))
    ref = o.fastconv64(x, w)
    err = o.rel_l2(y, ref)
    say("  rel_l2", err, "db", o.diff_db(y, ref))
    assert err < REL_L2
    assert o.diff_db(y, ref) < DB


def test_fastconv_fourstep_coefficient_matrix(ovx):
    rows, n = 3, 32768
    rng = np.random.default_rng(21)
    x, c = cg(rng, (rows, n)), cg(rng, (rows, n))
    fc = ovx.Fastconv(rows, n)
    y = host(fc(dev(x), dev(c)))
    assert o.rel_l2(y, o.fastconv64(x, c)) < REL_L2


# ---- kernel variants and the resident-replica path ----------------------------

@pytest.mark.parametrize("n", [1024, 2048, 4096, 8192, 16384])
def test_every_variant_matches_fp64(ovx, n):
    """Each compiled kernel variant of a length (the autotuner's choices) gives
    the same result to the stated tolerance: fastconv, forward and inverse."""
    rows = 21
    rng = np.random.default_rng(n + 5)
    x = cg(rng, (rows, n))
    w = cg(rng, n)
    c = cg(rng, (rows, n))
    ref = o.fastconv64(x, w)
    refc = o.fastconv64(x, c)
    reff = o.fftm64(x.astype(np.complex128), o.FWD)
    refi = o.fftm64(x.astype(np.complex128), o.INV, 1.0 / n)
    for v in range(ovx.variants(n)):
        say("variant", n, v)
        fc = ovx.Fastconv(rows, n, variant=v)
        assert o.rel_l2(host(fc(dev(x), dev(w))), ref) < REL_L2, (n, v, fc.name)
        assert o.rel_l2(host(fc(dev(x), dev(c))), refc) < REL_L2, (n, v, fc.name)
        ff = ovx.Fftm(rows, n, ovx.ROW, ovx.FFT_FWD, 1.0, variant=v)
        assert o.rel_l2(host(ff(dev(x))), reff) < REL_L2, (n, v, ff.name)
        fi = ovx.Fftm(rows, n, ovx.ROW, ovx.FFT_INV, 1.0 / n, variant=v)
        assert o.rel_l2(host(fi(dev(x))), refi) < REL_L2, (n, v, fi.name)


@pytest.mark.parametrize("n", [512, 2048, 3072, 16384, 65536])
def test_fastconv_resident_replica(ovx, n):
    rows = 9
    rng = np.random.default_rng(n + 7)
    x = cg(rng, (rows, n))
    chirp = wl.lfm_chirp(n // 4, n // 32).astype(np.complex64)
    rep = ovx.Replica(n, chirp)
    fc = ovx.Fastconv(rows, n)
    y = host(fc.with_replica(dev(x), rep))
    w = wl.replica_spectrum(chirp.astype(np.complex128), n)
    assert o.rel_l2(y, o.fastconv64(x, w)) < REL_L2


def test_autotune_records_and_plans_use_it(ovx, tmp_path, monkeypatch):
    monkeypatch.setenv("OVX_TUNE_CACHE", str(tmp_path / "tune.txt"))
    v, us = ovx.autotune(4096, 300)
    assert 0 <= v < ovx.variants(4096) and us > 0
    # rows 300 -> bucket 512; 400 shares it
    assert ovx.tuned_variant(4096, 400) == v
    assert (tmp_path / "tune.txt").exists()
    rng = np.random.default_rng(31)
    x = cg(rng, (300, 4096))
    w = cg(rng, 4096)
    fc = ovx.Fastconv(300, 4096)
    y = host(fc(dev(x), dev(w)))
    assert o.rel_l2(y, o.fastconv64(x, w)) < REL_L2


# the reference's own size list (tests/correlation.cpp:225-289: corr_cover and
# impl_corr_cover), each with its three input patterns (:68-82), at its
# threshold (-100 dB, :20-29)
CORR_COVER = [(8, 8), (1, 128), (7, 128), (8, 128), (9, 128), (7, 127), (8, 127), (9, 127), (7, 129),
              (8, 129), (9, 129), (32, 128), (64, 128), (96, 128), (12, 96), (12, 97), (11, 97), (12, 98)]


@pytest.mark.parametrize("support", ["full", "same", "min"])
@pytest.mark.parametrize("bias", ["biased", "unbiased"])
@pytest.mark.parametrize("m,n", CORR_COVER)
def test_correlation_reference_cases(ovx, support, bias, m, n):
    rng = np.random.default_rng(1000 + 7 * m + n)
    ramp = np.arange(n, dtype=np.float64).astype(np.complex64)
    c = ovx.Correlation(support, m, n)
    for ref, x in ((np.ones(m, np.complex64), ramp),
                   (cg(rng, m), ramp),
                   (cg(rng, m), cg(rng, n))):
        y = host(c(bias, dev(ref), dev(x)))
        exp = o.corr64(bias, support, ref, x)
        assert o.diff_db(exp, y) < DB, (support, bias, m, n)


def _conv_cover():
    """(symmetry, input size, coefficient count, decimation) of the
    reference's cases() (tests/convolution/convolution.hpp:453-489)."""
    c = [("nonsym", 8, 8, 1), ("nonsym", 5, 5, 1), ("sym_even_len_even", 8, 4, 1), ("sym_even_len_odd", 7, 4, 1)]
    c += [("nonsym", 5, 4, d) for d in (1, 2, 3, 4)]
    for size in (32, 128, 512):
        c += [("nonsym", size, 8, 1), ("nonsym", 2 * size, 7, 2), ("nonsym", size + 4, 6, 3),
              ("sym_even_len_even", size, 5, 1), ("sym_even_len_even", size + 1, 6, 2),
              ("sym_even_len_odd", size, 4, 1), ("sym_even_len_odd", size + 3, 3, 2)]
    return c


@pytest.mark.parametrize("support", ["full", "same", "min"])
@pytest.mark.parametrize("symmetry,n,m,d", _conv_cover())
def test_convolution_reference_cases(ovx, symmetry, n, m, d, support):
    rng = np.random.default_rng(n * 31 + m * 7 + d)
    coeff, x = cg(rng, m), cg(rng, n)
    conv = ovx.Convolution(coeff, symmetry, support, n, d)
    y = host(conv(dev(x)))
    exp = o.conv64(support, o.conv_kernel(symmetry, coeff), x, d)
    assert y.shape == exp.shape
    assert o.diff_db(exp, y) < DB, (symmetry, n, m, d, support)
