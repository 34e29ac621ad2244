"""Python mirror of the VSIPL++ objects on the fast-convolution path.

Each class wraps one C-ABI plan (include/ovxhip.h) and mirrors the reference
object of the same name (paths relative to the OpenVSIP tree):

  Fftm         vsip::Fftm            src/vsip/impl/signal/fft.hpp:134-161
  Fft          vsip::Fft             src/vsip/impl/signal/fft.hpp:75-103
  Fastconv     inv_fftm(vmmul<0>(w, fwd_fftm(x)))  tests/fastconv.cpp:181-213
  vmmul        vsip::vmmul<D>        src/vsip/impl/vmmul.hpp:50-128
  Correlation  vsip::Correlation     src/vsip/impl/signal/corr.hpp:55-122
  Convolution  vsip::Convolution     src/vsip/impl/signal/conv.hpp:55-108
  Storage      ovxx::storage_manager src/ovxx/storage/manager.hpp:207-360
                                     (host location + a HIP device location)

Data are torch complex64 tensors on the GPU (torch provides device memory and
streams only); work is enqueued on torch's current stream.  Every call goes
through libovxhip.so; there is no fallback path.
"""
import ctypes

from . import _lib
from ._lib import check, lib

FFT_FWD = _lib.FFT_FWD
FFT_INV = _lib.FFT_INV
ROW = _lib.ROW
COL = _lib.COL


def _torch():
    import torch
    return torch


def _stream():
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _check_tensor(t, name):
    torch = _torch()
    if t.dtype != torch.complex64:
        raise TypeError(f"{name} must be complex64 (complex<float>)")
    if not t.is_cuda:
        raise TypeError(f"{name} must be a GPU tensor")


class _Plan:
    def __init__(self):
        self._h = ctypes.c_void_p()

    def close(self):
        if self._h:
            lib().ovx_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def name(self):
        return lib().ovx_plan_name(self._h).decode()

    @property
    def device_bytes(self):
        return lib().ovx_plan_device_bytes(self._h)


def variants(n):
    """Number of compiled kernel variants for transform length n."""
    return lib().ovx_length_variants(n)


def autotune(n, rows):
    """Time every kernel variant of length n at `rows` rows; cache and return
    (best variant, microseconds)."""
    v = ctypes.c_int()
    us = ctypes.c_double()
    check(lib().ovx_autotune_fastconv(n, rows, ctypes.byref(v), ctypes.byref(us)))
    return v.value, us.value


def tuned_variant(n, rows):
    return lib().ovx_tuned_variant(n, rows)


def autotune_batch(n, rows):
    """Time the batch tile bodies of length n on a rows x n job; cache and
    return (best variant, microseconds per one-job batch replay)."""
    v = ctypes.c_int()
    us = ctypes.c_double()
    check(lib().ovx_autotune_batch(n, rows, ctypes.byref(v), ctypes.byref(us)))
    return v.value, us.value


def tuned_batch_variant(n, rows):
    return lib().ovx_tuned_batch_variant(n, rows)


def _flags(variant):
    return 0 if variant is None else _lib.FLAG_VARIANT(variant + 1)


class Fftm(_Plan):
    """Multiple FFTs along `axis` (ROW: each row) with in-kernel scaling."""

    def __init__(self, rows, cols, axis=ROW, direction=FFT_FWD, scale=1.0, variant=None):
        super().__init__()
        self.rows, self.cols, self.axis, self.direction, self.scale = rows, cols, axis, direction, scale
        check(lib().ovx_fftm_plan_create(ctypes.byref(self._h), rows, cols, axis, direction, scale,
                                         _flags(variant)))

    def __call__(self, x, out=None):
        torch = _torch()
        _check_tensor(x, "input")
        if out is None:
            out = torch.empty_like(x)
        _check_tensor(out, "output")
        check(lib().ovx_fftm_out_of_place(self._h, _ptr(x), x.stride(0), x.stride(1), _ptr(out),
                                          out.stride(0), out.stride(1), self.rows, self.cols, _stream()))
        return out

    def in_place(self, x):
        _check_tensor(x, "inout")
        check(lib().ovx_fftm_in_place(self._h, _ptr(x), x.stride(0), x.stride(1), self.rows, self.cols,
                                      _stream()))
        return x


class FftmReal(_Plan):
    """Real <-> complex Fftm over a rows x cols real domain.  FFT_FWD maps
    real float32 lines of even length n to n/2 + 1 complex values; FFT_INV
    maps n/2 + 1 complex values back to n real ones (unnormalised, times
    `scale`)."""

    def __init__(self, rows, cols, axis=ROW, direction=FFT_FWD, scale=1.0, variant=None):
        super().__init__()
        self.rows, self.cols, self.axis, self.direction, self.scale = rows, cols, axis, direction, scale
        check(lib().ovx_fftm_real_plan_create(ctypes.byref(self._h), rows, cols, axis, direction, scale,
                                              _flags(variant)))

    def complex_shape(self):
        if self.axis == ROW:
            return (self.rows, self.cols // 2 + 1)
        return (self.rows // 2 + 1, self.cols)

    def __call__(self, x, out=None):
        torch = _torch()
        if out is None:
            if self.direction == FFT_FWD:
                shape = self.complex_shape()
                dtype = torch.complex64
            else:
                shape = (self.rows, self.cols)
                dtype = torch.float32
            if self.axis == ROW:
                out = torch.empty(shape, dtype=dtype, device=x.device)
            else:  # column lines are contiguous: column-major storage
                out = torch.empty(shape[::-1], dtype=dtype, device=x.device).t()
        if not (x.is_cuda and out.is_cuda):
            raise ValueError("real Fftm tensors must live on the device")
        check(lib().ovx_fftm_real_out_of_place(self._h, _ptr(x), x.stride(0), x.stride(1), _ptr(out),
                                               out.stride(0), out.stride(1), self.rows, self.cols, _stream()))
        return out


class Fft(_Plan):
    def __init__(self, n, direction=FFT_FWD, scale=1.0):
        super().__init__()
        self.n = n
        check(lib().ovx_fft_plan_create(ctypes.byref(self._h), n, direction, scale, 0))

    def __call__(self, x, out=None):
        torch = _torch()
        _check_tensor(x, "input")
        if out is None:
            out = torch.empty_like(x)
        check(lib().ovx_fft_out_of_place(self._h, _ptr(x), x.stride(0), _ptr(out), out.stride(0), self.n,
                                         _stream()))
        return out


class Fastconv(_Plan):
    """Y = scale * IFFT_row( W (.) FFT_row(X) ) in one fused pass."""

    def __init__(self, rows, n, scale=None, variant=None, precision="fp32"):
        """precision: "fp32" (FP32 VALU stages, the parity path) or "bf16x3"
        (radix-16 stages on the MFMA cores as split BF16 products; n = 16384,
        one spectrum vector)."""
        super().__init__()
        self.rows, self.n = rows, n
        self.scale = 1.0 / n if scale is None else scale
        flags = _flags(variant)
        if precision == "bf16x3":
            flags |= _lib.FLAG_BF16X3
        elif precision != "fp32":
            raise ValueError("precision must be 'fp32' or 'bf16x3'")
        check(lib().ovx_fastconv_plan_create(ctypes.byref(self._h), rows, n, self.scale, flags))

    def __call__(self, x, w, out=None, rows=None):
        torch = _torch()
        _check_tensor(x, "input")
        _check_tensor(w, "spectrum")
        if out is None:
            out = torch.empty_like(x)
        rows = x.shape[0] if rows is None else rows
        ldw = 0 if w.dim() == 1 else w.stride(0)
        check(lib().ovx_fastconv_exec(self._h, _ptr(x), x.stride(0), _ptr(w), ldw, _ptr(out), out.stride(0),
                                      rows, _stream()))
        return out

    def with_replica(self, x, replica, out=None, rows=None):
        """Fastconv against a resident replica (spectrum prepared once)."""
        torch = _torch()
        _check_tensor(x, "input")
        if out is None:
            out = torch.empty_like(x)
        rows = x.shape[0] if rows is None else rows
        check(lib().ovx_fastconv_exec_replica(self._h, _ptr(x), x.stride(0), replica.id, _ptr(out),
                                              out.stride(0), rows, _stream()))
        return out


def vmmul(dim, v, m, out=None, alpha=1.0):
    """out = alpha * vmmul<dim>(v, m): dim 0 multiplies each row of m by v
    (len(v) == m.shape[1]), dim 1 each column (len(v) == m.shape[0]).  v and
    m are complex64 or float32 GPU tensors of any strides; out (complex64
    when either is complex, else float32) may be m."""
    torch = _torch()
    ok = (torch.complex64, torch.float32)
    for t, name in ((v, "v"), (m, "m")):
        if t.dtype not in ok or not t.is_cuda:
            raise TypeError(f"{name} must be a complex64 or float32 GPU tensor")
    if v.dim() != 1 or m.dim() != 2:
        raise ValueError("v must be 1-D and m 2-D")
    rows, cols = m.shape
    if v.shape[0] != (cols if dim == 0 else rows):
        raise ValueError(f"vmmul<{dim}>: vector length {v.shape[0]} does not match the matrix {rows}x{cols}")
    odt = torch.complex64 if torch.complex64 in (v.dtype, m.dtype) else torch.float32
    if out is None:
        out = torch.empty((rows, cols), dtype=odt, device=m.device)
    if out.dtype != odt or tuple(out.shape) != (rows, cols):
        raise TypeError("out has the wrong type or shape")
    check(lib().ovx_vmmul(dim, int(v.dtype == torch.complex64), _ptr(v), v.stride(0),
                          int(m.dtype == torch.complex64), _ptr(m), m.stride(0), m.stride(1), _ptr(out),
                          out.stride(0), out.stride(1), rows, cols, alpha, _stream()))
    return out


class Replica:
    """A replica spectrum kept resident on the device."""

    def __init__(self, n, replica, is_spectrum=False):
        import numpy as np
        r = np.ascontiguousarray(replica, dtype=np.complex64)
        self.n = n
        self._id = ctypes.c_int()
        check(lib().ovx_replica_create(n, r.ctypes.data_as(ctypes.c_void_p), r.size, int(bool(is_spectrum)),
                                       ctypes.byref(self._id)))

    @property
    def id(self):
        return self._id.value

    def spectrum_ptr(self):
        p = ctypes.c_void_p()
        n = ctypes.c_long()
        check(lib().ovx_replica_spectrum(self._id, ctypes.byref(p), ctypes.byref(n)))
        return p

    def spectrum(self):
        """A torch view (copy) of the resident spectrum."""
        torch = _torch()
        out = torch.empty(self.n, dtype=torch.complex64, device="cuda")
        torch.cuda.synchronize()
        import ctypes as C
        hip = _hip()
        hip.hipMemcpy(C.c_void_p(out.data_ptr()), self.spectrum_ptr(), C.c_size_t(self.n * 8), C.c_int(3))
        return out

    def close(self):
        if self._id.value:
            lib().ovx_replica_destroy(self._id)
            self._id = ctypes.c_int(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_hip_lib = None


class Storage:
    """A block's data in host memory and in HBM (ovx_storage): each location
    is valid or stale and a pointer request copies only into a stale one.
    `user` (a numpy array) becomes the host location when given; otherwise
    the library allocates pinned host memory on first host use.  Pointers
    are returned as ctypes.c_void_p for the C-ABI entry points."""

    HOST, DEVICE = _lib.LOC_HOST, _lib.LOC_DEVICE
    READ, WRITE, READ_WRITE = _lib.ACCESS_READ, _lib.ACCESS_WRITE, _lib.ACCESS_READ_WRITE

    def __init__(self, nbytes, user=None):
        self._h = ctypes.c_void_p()
        self._user = user  # the array must outlive the storage
        if user is not None and user.nbytes < nbytes:
            raise ValueError("user array smaller than the storage")
        check(lib().ovx_storage_create(ctypes.byref(self._h), nbytes,
                                       None if user is None else ctypes.c_void_p(user.ctypes.data)))

    def ptr(self, where, access, stream=None):
        p = ctypes.c_void_p()
        check(lib().ovx_storage_ptr(self._h, where, access, ctypes.byref(p), stream))
        return p

    def release(self, stream=None):
        check(lib().ovx_storage_release(self._h, stream))

    def admit(self, update=True):
        check(lib().ovx_storage_admit(self._h, int(bool(update))))

    def info(self):
        """(valid-location bits, bytes copied host->device, device->host)"""
        v, a, b = ctypes.c_uint(), ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().ovx_storage_info(self._h, ctypes.byref(v), ctypes.byref(a), ctypes.byref(b)))
        return v.value, a.value, b.value

    def close(self):
        if self._h:
            lib().ovx_storage_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _hip():
    global _hip_lib
    if _hip_lib is None:
        _hip_lib = ctypes.CDLL("libamdhip64.so")
        _hip_lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return _hip_lib


class Correlation(_Plan):
    def __init__(self, support, m, n):
        super().__init__()
        self.support, self.m, self.n = support, m, n
        check(lib().ovx_corr_plan_create(ctypes.byref(self._h), _lib.SUPPORT[support], m, n, 0))
        self.output_size = lib().ovx_corr_output_size(_lib.SUPPORT[support], m, n)

    def __call__(self, bias, ref, x, out=None):
        torch = _torch()
        _check_tensor(ref, "reference")
        _check_tensor(x, "input")
        if out is None:
            out = torch.empty(self.output_size, dtype=torch.complex64, device=x.device)
        check(lib().ovx_corr_exec(self._h, _lib.BIAS[bias], _ptr(ref), _ptr(x), _ptr(out), _stream()))
        return out


class Convolution(_Plan):
    def __init__(self, coeff, symmetry, support, n, decimation=1):
        super().__init__()
        import numpy as np
        c = np.ascontiguousarray(coeff, dtype=np.complex64)
        self.support, self.n, self.decimation = support, n, decimation
        check(lib().ovx_conv_plan_create(ctypes.byref(self._h), _lib.SYMMETRY[symmetry], _lib.SUPPORT[support],
                                         c.ctypes.data_as(ctypes.c_void_p), c.size, n, decimation, 0))
        m = {"nonsym": c.size, "sym_even_len_odd": 2 * c.size - 1, "sym_even_len_even": 2 * c.size}[symmetry]
        self.output_size = lib().ovx_conv_output_size(_lib.SUPPORT[support], m, n, decimation)

    def __call__(self, x, out=None):
        torch = _torch()
        _check_tensor(x, "input")
        if out is None:
            out = torch.empty(self.output_size, dtype=torch.complex64, device=x.device)
        check(lib().ovx_conv_exec(self._h, _ptr(x), _ptr(out), _stream()))
        return out


def init(device=0):
    check(lib().ovx_init(device))


class Batch:
    """Uneven fastconv jobs packed into persistent tile-queue launches
    (graph replay)."""

    def __init__(self):
        self._h = ctypes.c_void_p()
        check(lib().ovx_batch_create(ctypes.byref(self._h)))
        self._keep = []

    def add(self, x, replica, out):
        _check_tensor(x, "input")
        _check_tensor(out, "output")
        rows, n = x.shape
        check(lib().ovx_batch_add_fastconv(self._h, n, rows, _ptr(x), x.stride(0), replica.id, _ptr(out),
                                           out.stride(0)))
        self._keep.append((x, replica, out))

    def commit(self):
        check(lib().ovx_batch_commit(self._h, _stream()))

    def __call__(self):
        check(lib().ovx_batch_exec(self._h, _stream()))

    def plain(self):
        """One standalone launch per job (the per-request path)."""
        check(lib().ovx_batch_exec_plain(self._h, _stream()))

    @property
    def tiles(self):
        return lib().ovx_batch_tiles(self._h)

    @property
    def launches(self):
        return lib().ovx_batch_launches(self._h)

    def job_variants(self):
        """Kernel variant of every job (in the order added), as Fastconv's
        `variant` argument takes it."""
        return [lib().ovx_batch_job_variant(self._h, k) for k in range(len(self._keep))]

    def close(self):
        if self._h:
            lib().ovx_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CorrReference:
    """A correlation reference prepared once for serving (H = FFT(reversed ref))."""

    def __init__(self, support, ref, n):
        import numpy as np
        r = np.ascontiguousarray(ref, dtype=np.complex64)
        self.support, self.m, self.n = support, r.size, n
        self._id = ctypes.c_int()
        check(lib().ovx_corr_reference_create(_lib.SUPPORT[support], r.size, n, r.ctypes.data_as(ctypes.c_void_p),
                                              ctypes.byref(self._id)))
        self.output_size = lib().ovx_corr_output_size(_lib.SUPPORT[support], r.size, n)

    @property
    def id(self):
        return self._id.value

    def close(self):
        if self._id.value:
            lib().ovx_corr_reference_destroy(self._id)
            self._id = ctypes.c_int(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fastconv_request(x, replica, y, t_arrival=0.0):
    rows, n = x.shape
    return _lib.Request(t_arrival, _lib.REQ_FASTCONV, n, rows, x.data_ptr(), x.stride(0), y.data_ptr(), y.stride(0),
                        replica.id, 0)


def fftm_request(x, y, t_arrival=0.0):
    rows, n = x.shape
    return _lib.Request(t_arrival, _lib.REQ_FFTM, n, rows, x.data_ptr(), x.stride(0), y.data_ptr(), y.stride(0), 0, 0)


def corr_request(x, reference, y, bias="biased", t_arrival=0.0):
    return _lib.Request(t_arrival, _lib.REQ_CORR, reference.n, 1, x.data_ptr(), x.numel(), y.data_ptr(), y.numel(),
                        reference.id, _lib.BIAS[bias])


class Server:
    """Continuous-batching server (ovx_server_*); plain=True runs every request
    as its own standalone launch.  The server runs on its own streams, which
    do not wait for torch's: synchronize the stream that wrote a request's
    input (or is still writing its output) before submitting it."""

    def __init__(self, slots=3, max_jobs=512, max_tiles=65536, hbm_budget_mb=64 * 1024, workspace_mb=1024,
                 plain=False):
        cfg = _lib.ServerConfig(slots, 1 if plain else max_jobs, max_tiles, hbm_budget_mb, workspace_mb,
                                int(bool(plain)))
        self._h = ctypes.c_void_p()
        check(lib().ovx_server_create(ctypes.byref(self._h), ctypes.byref(cfg)))

    def submit(self, req):
        t = ctypes.c_ulong()
        check(lib().ovx_server_submit(self._h, ctypes.byref(req), ctypes.byref(t)))
        return t.value

    def wait(self, ticket, timeout=60.0):
        lat = ctypes.c_double()
        check(lib().ovx_server_wait(self._h, ticket, timeout, ctypes.byref(lat)))
        return lat.value

    def run_trace(self, reqs):
        """reqs: list of Request with t_arrival set; returns (latencies, makespan)."""
        import numpy as np
        arr = (_lib.Request * len(reqs))(*reqs)
        lat = np.zeros(len(reqs), dtype=np.float64)
        mk = ctypes.c_double()
        check(lib().ovx_server_run_trace(self._h, ctypes.cast(arr, ctypes.c_void_p), len(reqs),
                                         lat.ctypes.data_as(ctypes.c_void_p), ctypes.byref(mk)))
        return lat, mk.value

    def close(self):
        if self._h:
            lib().ovx_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
