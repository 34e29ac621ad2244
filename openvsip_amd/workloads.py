"""Seeded synthetic inputs for the fast-convolution workloads (numpy).

Shapes and seeds:
  pulse_compression_small   64 x 2048, seed 0, chirp length 2048, B*T = 512
  radar_256x16384           256 x 16384, seed 1, 5 chirp echoes per row,
                            chirp length 4096, B*T = 1024
  radar_64x262144           64 x 262144, same construction
  batch_jobs                512 jobs, seed 3, N from a fixed set, rows U[16, 256],
                            chirp B*T = N/8

Complex Gaussian entries have re, im ~ N(0, 1/2).  The linear-FM chirp of L
samples with time-bandwidth product BT is r[n] = exp(i pi BT (n/L)^2),
n < L (t = n/fs over a pulse of duration T = L/fs, so pi (B/T) t^2 =
pi BT (n/L)^2).  The replica spectrum is W = FFT_N(r zero-padded to N),
computed in FP64 and rounded to complex64.
"""
import numpy as np

BATCH_SIZES = (512, 1024, 2048, 4096, 8192, 16384, 65536, 3072, 6144, 12288)


def cgauss(rng, shape):
    s = np.sqrt(0.5)
    return (rng.standard_normal(shape) * s + 1j * rng.standard_normal(shape) * s).astype(np.complex64)


def lfm_chirp(length, bt):
    n = np.arange(length, dtype=np.float64)
    return np.exp(1j * np.pi * bt * (n / length) ** 2)


def replica_spectrum(chirp, n):
    pad = np.zeros(n, dtype=np.complex128)
    pad[: chirp.size] = chirp
    return np.fft.fft(pad).astype(np.complex64)


def pulse_compression_small(seed=0, rows=64, n=2048, bt=512):
    rng = np.random.default_rng(seed)
    x = cgauss(rng, (rows, n))
    w = replica_spectrum(lfm_chirp(n, bt), n)
    return x, w


def radar(rows, n, seed=1, chirp_len=4096, bt=1024, echoes=5):
    """Gaussian noise plus `echoes` chirp echoes per row at uniform delays in
    [0, n - chirp_len] with amplitudes U[0.1, 1.0].  Returns x, W, delays,
    amplitudes."""
    rng = np.random.default_rng(seed)
    x = cgauss(rng, (rows, n))
    chirp = lfm_chirp(chirp_len, bt).astype(np.complex64)
    delays = rng.integers(0, n - chirp_len + 1, size=(rows, echoes))
    amps = rng.uniform(0.1, 1.0, size=(rows, echoes)).astype(np.float32)
    for r in range(rows):
        for e in range(echoes):
            d = delays[r, e]
            x[r, d:d + chirp_len] += amps[r, e] * chirp
    w = replica_spectrum(chirp, n)
    return x, w, delays, amps


def batch_jobs(seed=3, count=512, sizes=BATCH_SIZES, min_rows=16, max_rows=256):
    """Job descriptors (n, rows, replica_bt) of the batched workload."""
    rng = np.random.default_rng(seed)
    ns = rng.choice(np.array(sizes), size=count)
    rows = rng.integers(min_rows, max_rows + 1, size=count)
    return [(int(n), int(r)) for n, r in zip(ns, rows)]


def batch_job_data(n, rows, seed):
    rng = np.random.default_rng(seed)
    x = cgauss(rng, (rows, n))
    w = replica_spectrum(lfm_chirp(n, n // 8), n)
    return x, w
