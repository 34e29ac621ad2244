"""Batched uneven jobs in persistent launches: each job's result is the
same alone and in the mix (bitwise: the batch runs the same device bodies as
the standalone kernels), and matches the FP64 oracle to rel-L2 <= 1e-5."""
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


def make_jobs(ovx, spec, seed=0):
    import torch
    jobs = []
    reps = {}
    for i, (n, rows) in enumerate(spec):
        x, w = wl.batch_job_data(n, rows, seed + i)
        if n not in reps:
            reps[n] = (ovx.Replica(n, w, is_spectrum=True), w)
        xd = dev(x)
        yd = torch.full_like(xd, float("nan"))
        jobs.append((n, rows, x, xd, yd))
    return jobs, reps


SPEC = [(512, 37), (1024, 16), (2048, 50), (3072, 19), (4096, 23), (6144, 17), (8192, 40),
        (12288, 21), (16384, 33), (65536, 16), (2048, 3), (512, 256)]


def test_batch_matches_oracle_and_standalone(ovx):
    jobs, reps = make_jobs(ovx, SPEC)
    b = ovx.Batch()
    for n, rows, x, xd, yd in jobs:
        b.add(xd, reps[n][0], yd)
    b.commit()
    print("tiles", b.tiles, "launches", b.launches, flush=True)
    # two persistent launches (1024-thread workgroups for the rows of 12288
    # and 16384 points, 512-thread ones for the shorter rows) + one launch
    # per four-step pass (line length 256)
    assert b.launches == 5
    b()
    outs = [host(yd) for *_, yd in jobs]
    variants = b.job_variants()
    for (n, rows, x, xd, yd), y, variant in zip(jobs, outs, variants):
        print("job", n, rows, "variant", variant, flush=True)
        assert np.isfinite(y).all(), (n, rows)
        err = o.rel_l2(y, o.fastconv64(x, reps[n][1]))
        assert err < 1e-5, (n, rows, err)
        # the tile body of every job is the device body of one standalone
        # variant (the autotuner's batch choice for its (n, rows))
        alone = ovx.Fastconv(rows, n, variant=variant).with_replica(xd, reps[n][0])
        assert np.array_equal(host(alone), y), (n, rows)
    # replay: same results again (graph replays are independent)
    for *_, yd in jobs:
        yd.fill_(float("nan"))
    b()
    b()
    for (n, rows, x, xd, yd), y in zip(jobs, outs):
        assert np.array_equal(host(yd), y), (n, rows)
    # the plain one-launch-per-job path gives the same bits
    for *_, yd in jobs:
        yd.fill_(float("nan"))
    b.plain()
    for (n, rows, x, xd, yd), y in zip(jobs, outs):
        assert np.array_equal(host(yd), y), (n, rows)


@pytest.mark.parametrize("spec,launches", [
    ([(1024, 5), (16384, 2), (512, 70)], 2),  # both persistent launches
    ([(1024, 5), (8192, 3), (512, 70)], 1),   # 512-thread workgroups only
    ([(12288, 3), (16384, 2)], 1),            # 1024-thread workgroups only
])
def test_batch_lds_only_persistent_launches(ovx, spec, launches):
    jobs, reps = make_jobs(ovx, spec, seed=40)
    b = ovx.Batch()
    for n, rows, x, xd, yd in jobs:
        b.add(xd, reps[n][0], yd)
    b.commit()
    assert b.launches == launches
    b()
    outs = [host(yd) for *_, yd in jobs]
    for (n, rows, x, xd, yd), y in zip(jobs, outs):
        assert o.rel_l2(y, o.fastconv64(x, reps[n][1])) < 1e-5, (n, rows)
    b.plain()
    for (n, rows, x, xd, yd), y in zip(jobs, outs):
        assert np.array_equal(host(yd), y), (n, rows)


def test_batch_four_step_only_and_mixed_line_lengths(ovx):
    """Four-step jobs alone (no persistent launch), including 131072 = 256 x
    512, whose middle pass runs lines of 512: one launch per non-empty
    (pass, line length) queue."""
    jobs, reps = make_jobs(ovx, [(65536, 3), (131072, 2)], seed=41)
    b = ovx.Batch()
    for n, rows, x, xd, yd in jobs:
        b.add(xd, reps[n][0], yd)
    b.commit()
    assert b.launches == 4  # pass 1 (256), pass 2 (256 and 512), pass 3 (256)
    b()
    for n, rows, x, xd, yd in jobs:
        y = host(yd)
        assert o.rel_l2(y, o.fastconv64(x, reps[n][1])) < 1e-5, n
        alone = ovx.Fastconv(rows, n).with_replica(xd, reps[n][0])
        assert np.array_equal(host(alone), y), n


def test_batch_large_four_step_jobs_bitwise_equal_standalone(ovx):
    """Jobs of 65536 with 300 and 100 rows (150 + 50 MiB) and a small
    LDS-resident job.  The four-step rows run in chunks of 128 MiB (256 rows
    of 65536), the first job's rows split over both chunks, each chunk's
    three passes back to back: one persistent launch plus 3 per chunk.
    Results are bitwise those of the standalone kernels."""
    import torch
    rng = np.random.default_rng(42)
    jobs = []
    reps = {}
    for n, rows in [(65536, 300), (65536, 100), (1024, 8)]:
        if n not in reps:
            reps[n] = ovx.Replica(n, wl.cgauss(rng, n), is_spectrum=True)
        xd = torch.randn(rows, n, dtype=torch.complex64, device="cuda")
        yd = torch.full_like(xd, float("nan"))
        jobs.append((n, rows, xd, yd))
    b = ovx.Batch()
    for n, rows, xd, yd in jobs:
        b.add(xd, reps[n], yd)
    b.commit()
    assert b.launches == 1 + 3 * 2
    b()
    for n, rows, xd, yd in jobs:
        alone = ovx.Fastconv(rows, n, variant=0).with_replica(xd, reps[n])
        assert np.array_equal(host(alone), host(yd)), (n, rows)
