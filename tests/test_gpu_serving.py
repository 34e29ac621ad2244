"""Continuous-batching server: a mixed stream of fastconv, forward Fftm and
correlation requests.  Every result is bitwise the same through the batched
server and through the plain one-launch-per-request server (the same device
bodies run in both), and matches the FP64 oracle to rel-L2 <= 1e-5."""
import numpy as np
import pytest

import oracle as o
from openvsip_amd import workloads as wl

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def build_mix(ovx):
    import torch
    rng = np.random.default_rng(11)
    specs = [("fc", 512, 40), ("fc", 2048, 17), ("fc", 3072, 9), ("fc", 16384, 5), ("fc", 65536, 3),
             ("fc", 12288, 6), ("fft", 1024, 12), ("fft", 6144, 7), ("fft", 65536, 2), ("fft", 512, 33),
             ("corr", 65536, 1), ("fc", 8192, 30), ("fft", 16384, 4), ("corr", 65536, 1), ("fc", 4096, 64),
             # every LDS-resident length of both kinds, in both persistent
             # launches (512-thread workgroups up to 8192, 1024 above)
             ("fc", 1024, 21), ("fc", 6144, 11), ("fft", 2048, 5), ("fft", 3072, 6), ("fft", 4096, 8),
             ("fft", 8192, 3), ("fft", 12288, 2)]
    reps, items = {}, []
    ref = wl.cgauss(rng, 1024)
    cref = ovx.CorrReference("full", ref, 65536)
    for i, (kind, n, rows) in enumerate(specs):
        if kind == "corr":
            x = wl.cgauss(rng, n)
            items.append(dict(kind=kind, n=n, x=x, xd=dev(x), bias="unbiased" if i % 2 else "biased",
                              ref=ref, cref=cref))
            continue
        x = wl.cgauss(rng, (rows, n))
        it = dict(kind=kind, n=n, rows=rows, x=x, xd=dev(x))
        if kind == "fc":
            if n not in reps:
                w = wl.replica_spectrum(wl.lfm_chirp(n, n // 8), n)
                reps[n] = (ovx.Replica(n, w, is_spectrum=True), w)
            it["rep"] = reps[n]
        items.append(it)
    return items, reps, cref


def run(ovx, items, plain):
    import torch
    srv = ovx.Server(slots=3, max_jobs=8, hbm_budget_mb=4096, workspace_mb=256, plain=plain)
    outs, reqs = [], []
    for it in items:
        if it["kind"] == "corr":
            y = torch.full((it["cref"].output_size,), float("nan"), dtype=torch.complex64, device="cuda")
            req = ovx.corr_request(it["xd"], it["cref"], y, it["bias"])
        else:
            y = torch.full_like(it["xd"], float("nan"))
            req = ovx.fastconv_request(it["xd"], it["rep"][0], y) if it["kind"] == "fc" else \
                ovx.fftm_request(it["xd"], y)
        reqs.append(req)
        outs.append(y)
    # the server's streams do not wait for torch's: the inputs and the NaN
    # fills of the outputs must have landed before the requests go in
    torch.cuda.synchronize()
    tickets = [srv.submit(req) for req in reqs]
    lats = [srv.wait(t) for t in tickets]
    res = [host(y) for y in outs]
    srv.close()
    return res, lats


def test_server_mix_same_alone_and_in_mix_and_vs_oracle(ovx):
    items, reps, cref = build_mix(ovx)
    batched, lat_b = run(ovx, items, plain=False)
    plain, lat_p = run(ovx, items, plain=True)
    for it, yb, yp in zip(items, batched, plain):
        print(it["kind"], it["n"], it.get("rows"), flush=True)
        assert np.isfinite(yb).all()
        assert np.array_equal(yb, yp), (it["kind"], it["n"])
        if it["kind"] == "fc":
            ref = o.fastconv64(it["x"], it["rep"][1])
        elif it["kind"] == "fft":
            ref = o.fftm64(it["x"].astype(np.complex128), o.FWD)
        else:
            ref = o.corr64(it["bias"], "full", it["ref"], it["x"])
        assert o.rel_l2(yb, ref) < 1e-5, (it["kind"], it["n"])
    assert all(l >= 0 for l in lat_b + lat_p)


def test_server_refuses_over_budget(ovx):
    import torch
    from openvsip_amd import OvxError
    srv = ovx.Server(slots=1, max_jobs=4, hbm_budget_mb=1, workspace_mb=16)
    x = torch.zeros((64, 4096), dtype=torch.complex64, device="cuda")  # 2 MiB + 2 MiB > 1 MiB
    y = torch.empty_like(x)
    w = wl.replica_spectrum(wl.lfm_chirp(4096, 512), 4096)
    rep = ovx.Replica(4096, w, is_spectrum=True)
    t = srv.submit(ovx.fastconv_request(x, rep, y))
    with pytest.raises(OvxError) as ei:
        srv.wait(t)
    assert ei.value.code == 5  # OVX_ERR_BUDGET
    srv.close()


def test_server_close_with_requests_in_flight(ovx):
    """Closing a server whose slots still run batches (and whose queue still
    holds requests) returns once the launched batches have retired; the
    device stays usable and every output a launched batch wrote is final."""
    import time
    import torch
    srv = ovx.Server(slots=3, max_jobs=2, hbm_budget_mb=4096, workspace_mb=64)
    w = wl.replica_spectrum(wl.lfm_chirp(16384, 2048), 16384)
    rep = ovx.Replica(16384, w, is_spectrum=True)
    xs = [torch.randn((64, 16384), dtype=torch.complex64, device="cuda") for _ in range(4)]
    ys = [torch.empty_like(x) for x in xs]
    torch.cuda.synchronize()  # the inputs are written before the server reads them
    for i in range(40):
        srv.submit(ovx.fastconv_request(xs[i % 4], rep, ys[i % 4]))
    t0 = time.time()
    srv.close()
    assert time.time() - t0 < 30.0
    torch.cuda.synchronize()
    z = torch.ones(4, device="cuda") * 2
    assert float(z.sum()) == 8.0


def test_server_large_four_step_batch(ovx):
    """A batch holding ~290 MB of four-step rows (three fastconv requests of
    65536 x 120 rows and a forward Fftm of 65536 x 100 rows, so fastconv and
    Fftm tiles share the pass queues) next to small LDS-resident requests:
    every result is bitwise the per-request server's and matches the FP64
    oracle on sampled rows."""
    import torch
    rng = np.random.default_rng(77)
    w = wl.replica_spectrum(wl.lfm_chirp(65536, 8192), 65536)
    rep = ovx.Replica(65536, w, is_spectrum=True)
    w2 = wl.replica_spectrum(wl.lfm_chirp(2048, 256), 2048)
    rep2 = ovx.Replica(2048, w2, is_spectrum=True)
    items = [("fc", 120), ("fc", 120), ("fft", 100), ("fc", 120), ("small", 40), ("small", 64)]
    xs = []
    for kind, rows in items:
        n = 2048 if kind == "small" else 65536
        xs.append((kind, torch.from_numpy(wl.cgauss(rng, (rows, n))).cuda()))

    def run(plain):
        srv = ovx.Server(slots=3, max_jobs=64, hbm_budget_mb=16384, workspace_mb=256, plain=plain)
        # keep every slot busy so that the requests below queue up and the
        # next free slot takes them as one batch
        busy_x = torch.randn((256, 65536), dtype=torch.complex64, device="cuda")
        busy_y = [torch.empty_like(busy_x) for _ in range(3)]
        outs, reqs = [], []
        for kind, x in xs:
            y = torch.full_like(x, float("nan"))
            if kind == "fc":
                req = ovx.fastconv_request(x, rep, y)
            elif kind == "fft":
                req = ovx.fftm_request(x, y)
            else:
                req = ovx.fastconv_request(x, rep2, y)
            reqs.append(req)
            outs.append(y)
        # the server's streams do not wait for torch's: every input and NaN
        # fill must have landed before the requests go in
        torch.cuda.synchronize()
        busy = [srv.submit(ovx.fastconv_request(busy_x, rep, y)) for y in busy_y]
        tickets = [srv.submit(req) for req in reqs]
        for t in busy + tickets:
            srv.wait(t)
        res = [host(y) for y in outs]
        srv.close()
        return res

    batched = run(False)
    plain = run(True)
    for (kind, x), yb, yp in zip(xs, batched, plain):
        assert np.isfinite(yb).all(), kind
        assert np.array_equal(yb, yp), kind
        xh = host(x)[:4]
        if kind == "fc":
            ref = o.fastconv64(xh, w)
        elif kind == "fft":
            ref = o.fftm64(xh.astype(np.complex128), o.FWD)
        else:
            ref = o.fastconv64(xh, w2)
        assert o.rel_l2(yb[:4], ref) < 1e-5, kind
