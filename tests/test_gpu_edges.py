"""Edge cases through the C-ABI on the GPU: empty inputs (zero rows) are a
no-op that succeeds and writes nothing, and the longest supported transform
(2^22, four-step with 2048 x 2048) matches the FP64 oracle."""
import numpy as np
import pytest

import oracle as o
from openvsip_amd import workloads as wl

pytestmark = pytest.mark.gpu

REL_L2 = 1e-5


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def test_zero_rows_is_a_noop(ovx):
    import torch
    n = 1024
    x = torch.ones((4, n), dtype=torch.complex64, device="cuda")
    y = torch.full((4, n), 7.0 + 0j, dtype=torch.complex64, device="cuda")
    w = torch.ones(n, dtype=torch.complex64, device="cuda")
    ovx.Fftm(0, n)(x[:0], out=y[:0])
    ovx.Fastconv(0, n)(x[:0], w, out=y[:0], rows=0)
    fc = ovx.Fastconv(4, n)
    fc(x, w, out=y, rows=0)                       # a plan for 4 rows run on none
    assert np.all(host(y) == 7.0)
    xr = torch.ones((0, n), dtype=torch.float32, device="cuda")
    yr = torch.empty((0, n // 2 + 1), dtype=torch.complex64, device="cuda")
    ovx.FftmReal(0, n)(xr, out=yr)


def test_longest_supported_length_fastconv_and_fftm(ovx):
    import torch
    n = int(ovx._lib.lib().ovx_max_length())
    assert n == 1 << 22
    rng = np.random.default_rng(22)
    x = wl.cgauss(rng, (1, n))
    w = wl.cgauss(rng, n)
    fc = ovx.Fastconv(1, n)
    y = host(fc(torch.from_numpy(x).cuda(), torch.from_numpy(w).cuda()))
    err = o.rel_l2(y, o.fastconv64(x, w))
    print(f"fastconv n={n} plan={fc.name} rel_l2={err:.3e}", flush=True)
    assert err < REL_L2
    f = ovx.Fftm(1, n, ovx.ROW, ovx.FFT_INV, 1.0 / n)
    z = host(f(torch.from_numpy(x).cuda()))
    assert o.rel_l2(z, o.fftm64(x, o.INV, 1.0 / n)) < REL_L2
