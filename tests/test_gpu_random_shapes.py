"""Seeded random draws over the transform and fastconv entry points: length
(every supported length up to 131072), row count, kernel variant, row
padding, axis, direction, in place or not, interleaved or split storage.
Each draw is checked against the FP64 oracle (rel-L2 <= 1e-5, the FP32 bar)
and, where rows are padded, for padding left as it was.  The draws are fixed
by the seed, so a failing case keeps its id."""
import ctypes

import numpy as np
import pytest

import oracle as o
from openvsip_amd import _lib
from openvsip_amd import workloads as wl

pytestmark = pytest.mark.gpu

REL_L2 = 1e-5
LENGTHS = [2, 3, 4, 5, 6, 8, 9, 10, 12, 16, 18, 20, 24, 32, 36, 40, 48, 64, 72, 80, 96, 128, 144, 160, 192, 256,
           288, 320, 384, 512, 576, 640, 768, 1024, 1152, 1280, 1536, 2048, 2304, 2560, 3072, 4096, 4608, 5120,
           6144, 8192, 9216, 10240, 12288, 16384, 32768, 65536, 131072]
OPS = ["fftm_row", "fftm_col", "fastconv_vec", "fastconv_mat", "fftm_split", "fastconv_split", "real"]
FILL = complex(5, -5)


def draw_cases(count=60, seed=2026):
    rng = np.random.default_rng(seed)
    cases = []
    for i in range(count):
        op = OPS[i % len(OPS)]
        n = int(rng.choice(LENGTHS))
        if op == "real":
            n = max(4, 2 * n if n % 2 else n)   # even, and n/2 a supported length
        rows = int(max(1, min(int(rng.integers(1, 40)), (1 << 21) // n)))
        cases.append(dict(op=op, n=n, rows=rows, gap=int(rng.choice([0, 1, 2, 17])),
                          inplace=bool(rng.integers(2)), inverse=bool(rng.integers(2)),
                          vdraw=int(rng.integers(8)), seed=int(rng.integers(1 << 30))))
    return cases


CASES = draw_cases()


def cid(c):
    return "{op}-n{n}-r{rows}-g{gap}-{ip}-{d}".format(ip="ip" if c["inplace"] else "oop",
                                                         d="inv" if c["inverse"] else "fwd", **c)


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def P(t):
    return ctypes.c_void_p(t.data_ptr())


@pytest.mark.parametrize("c", CASES, ids=[cid(c) for c in CASES])
def test_random_shape(ovx, c):
    import torch
    print("case", cid(c), flush=True)
    op, n, rows, gap = c["op"], c["n"], c["rows"], c["gap"]
    rng = np.random.default_rng(c["seed"])
    nv = ovx.variants(n)
    variant = c["vdraw"] % nv if nv > 0 else None
    direction = ovx.FFT_INV if c["inverse"] else ovx.FFT_FWD
    scale = 1.0 / n if c["inverse"] else 1.0
    odir = o.INV if c["inverse"] else o.FWD
    ld = n + gap
    lib = _lib.lib()

    if op == "fftm_row":
        x = wl.cgauss(rng, (rows, ld))
        f = ovx.Fftm(rows, n, ovx.ROW, direction, scale, variant=variant)
        ref = o.fftm64(x[:, :n].astype(np.complex128), odir, scale)
        if c["inplace"]:
            buf = torch.from_numpy(x).cuda()
            f.in_place(buf[:, :n])
        else:
            buf = torch.full((rows, ld), FILL, dtype=torch.complex64, device="cuda")
            f(torch.from_numpy(x).cuda()[:, :n], buf[:, :n])
        y = host(buf)
        assert o.rel_l2(y[:, :n], ref) < REL_L2
        assert np.all(y[:, n:] == (x[:, n:] if c["inplace"] else FILL))
    elif op == "fftm_col":
        # an n x rows matrix stored column-major with column stride ld
        zt = wl.cgauss(rng, (rows, ld))
        f = ovx.Fftm(n, rows, ovx.COL, direction, scale, variant=variant)
        ref = o.fftm64(zt[:, :n].astype(np.complex128), odir, scale)
        src = torch.from_numpy(zt).cuda()
        if c["inplace"]:
            f.in_place(src.t()[:n, :])
            y = host(src)
        else:
            out = torch.full((rows, ld), FILL, dtype=torch.complex64, device="cuda")
            f(src.t()[:n, :], out.t()[:n, :])
            y = host(out)
        assert o.rel_l2(y[:, :n], ref) < REL_L2
        assert np.all(y[:, n:] == (zt[:, n:] if c["inplace"] else FILL))
    elif op in ("fastconv_vec", "fastconv_mat"):
        x = wl.cgauss(rng, (rows, ld))
        w = wl.replica_spectrum(wl.lfm_chirp(n, max(1, n // 8)), n).astype(np.complex64)
        fc = ovx.Fastconv(rows, n, variant=variant)
        xd = torch.from_numpy(x).cuda()
        if op == "fastconv_mat":
            wm = np.zeros((rows, n + 3), np.complex64)
            for k in range(rows):
                wm[k, :n] = np.roll(w, 7 * k)
            wd = torch.from_numpy(wm).cuda()[:, :n]
            ref = np.stack([o.fastconv64(x[k:k + 1, :n], wm[k, :n])[0] for k in range(rows)])
        else:
            wd = torch.from_numpy(w).cuda()
            ref = o.fastconv64(x[:, :n], w)
        if c["inplace"]:
            fc(xd[:, :n], wd, out=xd[:, :n])
            y = host(xd)
        else:
            out = torch.full((rows, ld), FILL, dtype=torch.complex64, device="cuda")
            fc(xd[:, :n], wd, out=out[:, :n])
            y = host(out)
        assert o.rel_l2(y[:, :n], ref) < REL_L2
        assert np.all(y[:, n:] == (x[:, n:] if c["inplace"] else FILL))
    elif op == "fftm_split":
        z = wl.cgauss(rng, (rows, ld))
        f = ovx.Fftm(rows, n, ovx.ROW, direction, scale, variant=variant)
        re = torch.from_numpy(np.ascontiguousarray(z.real)).cuda()
        im = torch.from_numpy(np.ascontiguousarray(z.imag)).cuda()
        if c["inplace"]:
            _lib.check(lib.ovx_fftm_split_in_place(f._h, P(re), P(im), ld, 1, rows, n, None))
            yr, yi = re, im
        else:
            yr = torch.full((rows, ld), 5.0, device="cuda")
            yi = torch.full((rows, ld), -5.0, device="cuda")
            _lib.check(lib.ovx_fftm_split_out_of_place(f._h, P(re), P(im), ld, 1, P(yr), P(yi), ld, 1, rows, n,
                                                       None))
        y = host(yr) + 1j * host(yi)
        assert o.rel_l2(y[:, :n], o.fftm64(z[:, :n].astype(np.complex128), odir, scale)) < REL_L2
        assert np.all(y[:, n:] == (z[:, n:] if c["inplace"] else FILL))
    elif op == "fastconv_split":
        x = wl.cgauss(rng, (rows, ld))
        w = wl.replica_spectrum(wl.lfm_chirp(n, max(1, n // 8)), n).astype(np.complex64)
        fc = ovx.Fastconv(rows, n, variant=variant)
        xr = torch.from_numpy(np.ascontiguousarray(x.real)).cuda()
        xi = torch.from_numpy(np.ascontiguousarray(x.imag)).cuda()
        wr = torch.from_numpy(np.ascontiguousarray(w.real)).cuda()
        wi = torch.from_numpy(np.ascontiguousarray(w.imag)).cuda()
        if c["inplace"]:
            yr, yi = xr, xi
        else:
            yr = torch.full((rows, ld), 5.0, device="cuda")
            yi = torch.full((rows, ld), -5.0, device="cuda")
        _lib.check(lib.ovx_fastconv_split_exec(fc._h, P(xr), P(xi), ld, P(wr), P(wi), 0, P(yr), P(yi), ld, rows,
                                               None))
        y = host(yr) + 1j * host(yi)
        assert o.rel_l2(y[:, :n], o.fastconv64(x[:, :n], w)) < REL_L2
        assert np.all(y[:, n:] == (x[:, n:] if c["inplace"] else FILL))
    else:  # real: r2c forward, or c2r inverse with scale 1/n
        h = n // 2 + 1
        ldr = n + (gap & ~1)        # real line stride: even
        ldc = h + gap
        if not c["inverse"]:
            xr = rng.standard_normal((rows, ldr)).astype(np.float32)
            f = ovx.FftmReal(rows, n, ovx.ROW, ovx.FFT_FWD)
            out = torch.full((rows, ldc), FILL, dtype=torch.complex64, device="cuda")
            f(torch.from_numpy(xr).cuda()[:, :n], out=out[:, :h])
            y = host(out)
            assert o.rel_l2(y[:, :h], o.rfftm64(xr[:, :n])) < REL_L2
            assert np.all(y[:, h:] == FILL)
        else:
            xh = wl.cgauss(rng, (rows, ldc))
            f = ovx.FftmReal(rows, n, ovx.ROW, ovx.FFT_INV, scale=1.0 / n)
            out = torch.full((rows, ldr), 5.0, dtype=torch.float32, device="cuda")
            f(torch.from_numpy(xh).cuda()[:, :h], out=out[:, :n])
            y = host(out)
            assert o.rel_l2(y[:, :n].astype(np.complex128), o.irfftm64(xh[:, :h], n, scale=1.0 / n)) < REL_L2
            assert np.all(y[:, n:] == 5.0)
