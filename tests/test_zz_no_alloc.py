"""ovx_fastconv_exec_noalloc (the fused fastconv with its scratch in the
library's fixed device array instead of a plan buffer) against
ovx_fastconv_exec on the same plans and data: bitwise equal, on two calls in
a row, the second with a smaller input (fewer rows, or a shorter transform
on another plan), on the kernel paths that use scratch (hierarchical and
four-step plans permute a single spectrum into it) and on the ones that use
none (register-Stockham, a coefficient matrix); calls needing more than
OVX_NOALLOC_SCRATCH_BYTES are refused with OVX_ERR_UNSUPPORTED."""
import ctypes

import numpy as np
import pytest

from openvsip_amd import workloads as wl

pytestmark = pytest.mark.gpu

NOALLOC_BYTES = 4 << 20  # OVX_NOALLOC_SCRATCH_BYTES in include/ovxhip.h


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.complex64)).cuda()


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _run(ovx, fn, plan, x, w, ldw, y, rows):
    import torch
    st = fn(plan._h, _p(x), x.stride(0), _p(w), ldw, _p(y), y.stride(0), rows,
            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    ovx._lib.check(st)


def _pair(ovx, n, rows, variant, matrix, seed, rows_call=None):
    """(noalloc result, exec result) for one plan and one input."""
    import torch
    lib = ovx._lib.lib()
    rng = np.random.default_rng(seed)
    x = _dev(wl.cgauss(rng, (rows, n)))
    w = _dev(wl.cgauss(rng, (rows, n) if matrix else n))
    ldw = w.stride(0) if matrix else 0
    fc = ovx.Fastconv(rows, n, variant=variant)
    rc = rows if rows_call is None else rows_call
    y0 = torch.full_like(x, float("nan"))
    y1 = torch.full_like(x, float("nan"))
    _run(ovx, lib.ovx_fastconv_exec_noalloc, fc, x, w, ldw, y0, rc)
    _run(ovx, lib.ovx_fastconv_exec, fc, x, w, ldw, y1, rc)
    torch.cuda.synchronize()
    return fc.name, y0[:rc], y1[:rc]


CASES = [
    # (n, rows, variant (None = default), coefficient matrix)
    (16384, 64, 0, False),      # hierarchical: permuted spectrum in scratch
    (8192, 40, 0, True),        # hierarchical, coefficient matrix: no scratch
    (65536, 8, None, False),    # four-step: permuted spectrum in scratch
    (3072, 50, None, False),    # register-Stockham: no scratch
    (2048, 33, 1, False),       # register-Stockham variant of a hierarchical length
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}x{c[1]}-v{c[2]}-{'mat' if c[3] else 'vec'}")
def test_noalloc_equals_exec_twice_second_smaller(ovx, case):
    import torch
    n, rows, variant, matrix = case
    print("noalloc case", n, rows, variant, matrix, flush=True)
    name, a0, b0 = _pair(ovx, n, rows, variant, matrix, seed=n + rows)
    assert bool(torch.equal(a0, b0)), name
    assert bool(torch.isfinite(a0.abs()).all())
    # second call: a smaller input on the same plan (fewer rows) ...
    name, a1, b1 = _pair(ovx, n, rows, variant, matrix, seed=n + rows + 1, rows_call=rows // 2 + 1)
    assert bool(torch.equal(a1, b1)), name
    # ... and a shorter transform on another plan of the same kind
    name, a2, b2 = _pair(ovx, n // 2, rows // 2 + 1, variant, matrix, seed=n + rows + 2)
    assert bool(torch.equal(a2, b2)), name


def test_noalloc_refuses_more_scratch_than_its_limit(ovx):
    import torch
    lib = ovx._lib.lib()
    n, rows = 65536, 16   # a four-step coefficient matrix: 16 x 65536 x 8 bytes = 8 MiB of permuted spectrum
    assert rows * n * 8 > NOALLOC_BYTES
    fc = ovx.Fastconv(rows, n)
    x = torch.zeros((rows, n), dtype=torch.complex64, device="cuda")
    st = lib.ovx_fastconv_exec_noalloc(fc._h, _p(x), n, _p(x), n, _p(x), n, rows,
                                       ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert st == 2, lib.ovx_last_error()  # OVX_ERR_UNSUPPORTED
    host = ovx.Fastconv(4, 1024)
    host_plan = ctypes.c_void_p()
    ovx._lib.check(lib.ovx_fastconv_plan_create(ctypes.byref(host_plan), 4, 1024, 1.0 / 1024, 1))
    try:
        st = lib.ovx_fastconv_exec_noalloc(host_plan, _p(x), 1024, _p(x), 0, _p(x), 1024, 4, None)
        assert st == 2
    finally:
        lib.ovx_plan_destroy(host_plan)
    del host
