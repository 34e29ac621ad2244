"""vmmul<D>(v, m) on the GPU through the C-ABI (ovx_vmmul), restating the
reference's tests/vmmul.cpp:

* test_vmmul (vmmul.cpp:28-71): m(r, c) = r*cols + c, v = ramp(0, 1);
  vmmul<0> gives c*(r*cols + c), vmmul<1> gives r*(r*cols + c), exactly
  (integers below 2^24 are exact in float), also with the vector scaled by
  2, the matrix by 3 and the result negated (alpha = 2, 3, -1 here); every
  combination of row-/column-major input and output, the four value-type
  pairs (float, complex<float>), at 5 x 7, 8 x 8192 and 40 x 50
  (vmmul.cpp:150-176);
* test_subview / test_subview_unaligned (vmmul.cpp:74-147): strided and
  offset subviews of a larger matrix, written into the same subview of a
  zero matrix, compared with the per-row (per-column) product at the
  reference's criterion test::diff < -100 dB (support/test/diff.hpp:62-96),
  and nothing outside the subview touched;
* a large random complex case against numpy's float64 product
  (rel-L2 <= 1e-6), and vmmul in place (out = m)."""
import numpy as np
import pytest

import oracle as o

pytestmark = pytest.mark.gpu


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _dev(a, order):
    """numpy 2-D array -> GPU tensor in row- or column-major layout."""
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a if order == "row" else a.T)).cuda()
    return t if order == "row" else t.t()


def _empty(rows, cols, dtype, order):
    import torch
    t = torch.zeros((rows, cols) if order == "row" else (cols, rows), dtype=dtype, device="cuda")
    return t if order == "row" else t.t()


TYPES = {"f": (np.float32, "float32"), "c": (np.complex64, "complex64")}


@pytest.mark.parametrize("dim", [0, 1])
@pytest.mark.parametrize("m_order", ["row", "col"])
@pytest.mark.parametrize("out_order", ["row", "col"])
@pytest.mark.parametrize("types", ["ff", "cc", "fc", "cf"])
@pytest.mark.parametrize("shape", [(5, 7), (8, 8192), (40, 50)])
def test_vmmul_exact_integer_products(ovx, dim, m_order, out_order, types, shape):
    import torch
    # the reference runs its large shapes as vmmul<0> row-major only
    # (vmmul.cpp:166-170); here every layout runs at every shape
    rows, cols = shape
    vt, mt = TYPES[types[0]][0], TYPES[types[1]][0]
    m = (np.arange(rows)[:, None] * cols + np.arange(cols)[None, :]).astype(mt)
    vlen = cols if dim == 0 else rows
    v = np.arange(vlen).astype(vt)
    odt = torch.complex64 if "c" in types else torch.float32
    vd = torch.from_numpy(v).cuda()
    md = _dev(m, m_order)
    idx = (np.arange(cols)[None, :] if dim == 0 else np.arange(rows)[:, None])
    base = idx * (np.arange(rows)[:, None] * cols + np.arange(cols)[None, :])
    for alpha in (1.0, 2.0, -1.0, 3.0):
        out = _empty(rows, cols, odt, out_order)
        ovx.vmmul(dim, vd, md, out=out, alpha=alpha)
        got = host(out)
        want = (alpha * base).astype(got.dtype)
        if alpha != 3.0 or base.max() < 2 ** 24 // 3:
            # one rounding of an exact product (none below 2^24): bitwise
            assert np.array_equal(got, want), (dim, m_order, out_order, types, alpha)
        else:
            # alpha applies to the rounded product: one more rounding than
            # the reference's vmmul(v, 3*m) (its test::equal is approximate)
            assert np.max(np.abs(got - want) / np.maximum(np.abs(want), 1)) <= 2.0 ** -23


@pytest.mark.parametrize("types", ["ff", "cc"])
@pytest.mark.parametrize("case", [
    ("row", 0, (slice(0, 8), slice(0, 32)), (8, 32)),          # Domain<2>(8, 32)
    ("row", 0, (slice(0, 8, 2), slice(0, 32)), (8, 32)),       # Domain<2>(Domain<1>(0, 2, 4), 32)
    ("col", 1, (slice(0, 32), slice(0, 8)), (32, 8)),          # col2_type, col
    ("col", 1, (slice(0, 32), slice(0, 8, 2)), (32, 8)),       # Domain<2>(32, Domain<1>(0, 2, 4))
    ("row", 0, (slice(0, 8), slice(0, 32)), (8, 33)),          # unaligned: align 0, gap 1
    ("row", 0, (slice(0, 8), slice(1, 33)), (8, 33)),          # unaligned: align 1, gap 1
])
def test_vmmul_subviews(ovx, types, case):
    import torch
    order, sd, sub, full = case
    dt = np.float32 if types == "ff" else np.complex64
    a = np.ones(full, dtype=dt)
    sub_a = a[sub]
    vlen = sub_a.shape[1] if sd == 0 else sub_a.shape[0]
    w = np.arange(1, vlen + 1).astype(dt)
    ref = np.zeros(full, dtype=dt)
    ref[sub] = (w[None, :] if sd == 0 else w[:, None]) * sub_a
    ad = _dev(a, order)
    zd = _empty(full[0], full[1], torch.float32 if types == "ff" else torch.complex64, order)
    ovx.vmmul(sd, torch.from_numpy(w).cuda(), ad[sub], out=zd[sub])
    z = host(zd)
    assert o.diff_db(ref, z) < -100
    mask = np.ones(full, dtype=bool)
    mask[sub] = False
    assert np.all(z[mask] == 0)


def test_vmmul_random_complex_and_in_place(ovx):
    import torch
    rng = np.random.default_rng(5)
    rows, cols = 256, 16384
    m = (rng.standard_normal((rows, cols)) + 1j * rng.standard_normal((rows, cols))).astype(np.complex64)
    v0 = (rng.standard_normal(cols) + 1j * rng.standard_normal(cols)).astype(np.complex64)
    v1 = (rng.standard_normal(rows) + 1j * rng.standard_normal(rows)).astype(np.complex64)
    md = torch.from_numpy(m).cuda()
    y0 = host(ovx.vmmul(0, torch.from_numpy(v0).cuda(), md))
    r0 = v0.astype(np.complex128)[None, :] * m.astype(np.complex128)
    assert o.rel_l2(y0, r0) < 1e-6
    y1 = host(ovx.vmmul(1, torch.from_numpy(v1).cuda(), md))
    r1 = v1.astype(np.complex128)[:, None] * m.astype(np.complex128)
    assert o.rel_l2(y1, r1) < 1e-6
    ovx.vmmul(0, torch.from_numpy(v0).cuda(), md, out=md)  # in place
    assert np.array_equal(host(md), y0)


def test_vmmul_rejects_mismatched_sizes(ovx):
    import torch
    m = torch.zeros((4, 8), dtype=torch.complex64, device="cuda")
    with pytest.raises(ValueError):
        ovx.vmmul(0, torch.zeros(4, dtype=torch.complex64, device="cuda"), m)
    with pytest.raises(ovx.OvxInvalid):
        ovx._lib.check(ovx._lib.lib().ovx_vmmul(2, 1, None, 1, 1, None, 8, 1, None, 8, 1, 4, 8, 1.0, None))
