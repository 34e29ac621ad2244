"""A plan's device scratch and the calls that share it.

ovx_fastconv_exec on a hierarchical plan permutes the spectrum into a buffer
the plan owns, so a second call on another stream while the first may still
run would race on it: the C-ABI refuses it with OVX_ERR_STATE.  Calls that
use no plan scratch -- a resident replica's prepared spectrum -- are not
checked and run on any stream.  The first stream is held busy by a bounded
spin kernel queued ahead of the call."""
import ctypes

import numpy as np
import pytest

import oracle as o
from openvsip_amd import _lib
from openvsip_amd import workloads as wl

pytestmark = pytest.mark.gpu

R, N = 32, 16384
SPIN_CYCLES = 20_000_000  # a bounded spin on the GPU, milliseconds


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _s(stream):
    return ctypes.c_void_p(stream.cuda_stream)


def test_scratch_plan_refuses_a_second_stream_then_accepts_it():
    import torch
    import openvsip_amd as ovx

    ovx.init(0)
    rng = np.random.default_rng(11)
    x_np, w_np = wl.cgauss(rng, (R, N)), wl.cgauss(rng, (N,))
    x = torch.from_numpy(x_np).cuda()
    w = torch.from_numpy(w_np).cuda()
    ya, yb = torch.empty_like(x), torch.empty_like(x)
    fc = ovx.Fastconv(R, N, variant=0)
    assert "hierarchical" in fc.name
    lib = _lib.lib()
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    with torch.cuda.stream(sa):
        torch.cuda._sleep(SPIN_CYCLES)
    st = lib.ovx_fastconv_exec(fc._h, _p(x), N, _p(w), 0, _p(ya), N, R, _s(sa))
    assert st == _lib.OVX_OK
    st = lib.ovx_fastconv_exec(fc._h, _p(x), N, _p(w), 0, _p(yb), N, R, _s(sb))
    assert st == _lib.OVX_ERR_STATE
    torch.cuda.synchronize()
    st = lib.ovx_fastconv_exec(fc._h, _p(x), N, _p(w), 0, _p(yb), N, R, _s(sb))
    assert st == _lib.OVX_OK
    torch.cuda.synchronize()
    assert torch.equal(ya, yb)
    ref = o.fastconv64(x_np[:4], w_np)
    assert o.rel_l2(ya[:4].cpu().numpy(), ref) < 1e-5


def test_prepared_replica_calls_run_on_any_stream():
    import torch
    import openvsip_amd as ovx

    ovx.init(0)
    rng = np.random.default_rng(12)
    x_np, w_np = wl.cgauss(rng, (R, N)), wl.cgauss(rng, (N,))
    x = torch.from_numpy(x_np).cuda()
    rep = ovx.Replica(N, w_np, is_spectrum=True)
    ya, yb = torch.empty_like(x), torch.empty_like(x)
    fc = ovx.Fastconv(R, N, variant=0)
    lib = _lib.lib()
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    with torch.cuda.stream(sa):
        torch.cuda._sleep(SPIN_CYCLES)
    assert lib.ovx_fastconv_exec_replica(fc._h, _p(x), N, rep.id, _p(ya), N, R, _s(sa)) == _lib.OVX_OK
    assert lib.ovx_fastconv_exec_replica(fc._h, _p(x), N, rep.id, _p(yb), N, R, _s(sb)) == _lib.OVX_OK
    torch.cuda.synchronize()
    assert torch.equal(ya, yb)
    ref = o.fastconv64(x_np[:4], w_np)
    assert o.rel_l2(ya[:4].cpu().numpy(), ref) < 1e-5
