"""ctypes binding of libovxhip.so (the C-ABI declared in include/ovxhip.h).

The library is required: importing a symbol that needs it raises
OvxLibraryMissing when libovxhip.so has not been built -- there is no CPU or
PyTorch fallback for any operation.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libovxhip.so")

OVX_OK = 0
OVX_ERR_INVALID = 1
OVX_ERR_UNSUPPORTED = 2
OVX_ERR_NOMEM = 3
OVX_ERR_DEVICE = 4
OVX_ERR_BUDGET = 5
OVX_ERR_TIMEOUT = 6
OVX_ERR_STATE = 7

FFT_FWD = -2
FFT_INV = -1
ROW = 0
COL = 1
FLAG_HOST_POINTERS = 0x1
FLAG_BF16X3 = 0x2
LOC_HOST = 0
LOC_DEVICE = 1
ACCESS_READ = 1
ACCESS_WRITE = 2
ACCESS_READ_WRITE = 3


def FLAG_VARIANT(v):
    return (v & 0xFF) << 8

SUPPORT = {"full": 0, "same": 1, "min": 2}
BIAS = {"biased": 0, "unbiased": 1}
SYMMETRY = {"nonsym": 0, "sym_even_len_odd": 1, "sym_even_len_even": 2}

# every symbol include/ovxhip.h declares, with (restype, argtypes)
_P = ctypes.c_void_p
_L = ctypes.c_long
_I = ctypes.c_int
_U = ctypes.c_uint
_F = ctypes.c_float
_PP = ctypes.POINTER(ctypes.c_void_p)
SIGNATURES = {
    "ovx_init": (_I, [_I]),
    "ovx_last_error": (ctypes.c_char_p, []),
    "ovx_version": (ctypes.c_char_p, []),
    "ovx_max_length": (_L, []),
    "ovx_supported_length": (_I, [_L]),
    "ovx_length_variants": (_I, [_L]),
    "ovx_autotune_fastconv": (_I, [_L, _L, ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_double)]),
    "ovx_tuned_variant": (_I, [_L, _L]),
    "ovx_autotune_batch": (_I, [_L, _L, ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_double)]),
    "ovx_tuned_batch_variant": (_I, [_L, _L]),
    "ovx_fftm_plan_create": (_I, [_PP, _L, _L, _I, _I, _F, _U]),
    "ovx_fftm_out_of_place": (_I, [_P, _P, _L, _L, _P, _L, _L, _L, _L, _P]),
    "ovx_fftm_in_place": (_I, [_P, _P, _L, _L, _L, _L, _P]),
    "ovx_fftm_real_plan_create": (_I, [_PP, _L, _L, _I, _I, _F, _U]),
    "ovx_fftm_real_out_of_place": (_I, [_P, _P, _L, _L, _P, _L, _L, _L, _L, _P]),
    "ovx_fft_plan_create": (_I, [_PP, _L, _I, _F, _U]),
    "ovx_fft_out_of_place": (_I, [_P, _P, _L, _P, _L, _L, _P]),
    "ovx_fft_in_place": (_I, [_P, _P, _L, _L, _P]),
    "ovx_fastconv_plan_create": (_I, [_PP, _L, _L, _F, _U]),
    "ovx_fastconv_exec": (_I, [_P, _P, _L, _P, _L, _P, _L, _L, _P]),
    "ovx_fastconv_exec_noalloc": (_I, [_P, _P, _L, _P, _L, _P, _L, _L, _P]),
    "ovx_vmmul": (_I, [_I, _I, _P, _L, _I, _P, _L, _L, _P, _L, _L, _L, _L, _F, _P]),
    "ovx_replica_create": (_I, [_L, _P, _L, _I, ctypes.POINTER(_I)]),
    "ovx_replica_destroy": (_I, [_I]),
    "ovx_replica_spectrum": (_I, [_I, _PP, ctypes.POINTER(_L)]),
    "ovx_fastconv_exec_replica": (_I, [_P, _P, _L, _I, _P, _L, _L, _P]),
    "ovx_fftm_split_out_of_place": (_I, [_P, _P, _P, _L, _L, _P, _P, _L, _L, _L, _L, _P]),
    "ovx_fftm_split_in_place": (_I, [_P, _P, _P, _L, _L, _L, _L, _P]),
    "ovx_fft_split_out_of_place": (_I, [_P, _P, _P, _L, _P, _P, _L, _L, _P]),
    "ovx_fft_split_in_place": (_I, [_P, _P, _P, _L, _L, _P]),
    "ovx_fastconv_split_exec": (_I, [_P, _P, _P, _L, _P, _P, _L, _P, _P, _L, _L, _P]),
    "ovx_fftm_r2c_split": (_I, [_P, _P, _L, _L, _P, _P, _L, _L, _L, _L, _P]),
    "ovx_fftm_c2r_split": (_I, [_P, _P, _P, _L, _L, _P, _L, _L, _L, _L, _P]),
    "ovx_storage_create": (_I, [_PP, ctypes.c_size_t, _P]),
    "ovx_storage_destroy": (_I, [_P]),
    "ovx_storage_ptr": (_I, [_P, _I, _I, _PP, _P]),
    "ovx_storage_release": (_I, [_P, _P]),
    "ovx_storage_admit": (_I, [_P, _I]),
    "ovx_storage_info": (_I, [_P, ctypes.POINTER(_U), ctypes.POINTER(ctypes.c_uint64),
                              ctypes.POINTER(ctypes.c_uint64)]),
    "ovx_batch_create": (_I, [_PP]),
    "ovx_batch_add_fastconv": (_I, [_P, _L, _L, _P, _L, _I, _P, _L]),
    "ovx_batch_commit": (_I, [_P, _P]),
    "ovx_batch_exec": (_I, [_P, _P]),
    "ovx_batch_exec_plain": (_I, [_P, _P]),
    "ovx_batch_tiles": (_L, [_P]),
    "ovx_batch_launches": (_I, [_P]),
    "ovx_batch_job_variant": (_I, [_P, _L]),
    "ovx_batch_destroy": (_I, [_P]),
    "ovx_server_create": (_I, [_PP, ctypes.c_void_p]),
    "ovx_server_submit": (_I, [_P, ctypes.c_void_p, ctypes.POINTER(ctypes.c_ulong)]),
    "ovx_server_wait": (_I, [_P, ctypes.c_ulong, ctypes.c_double, ctypes.POINTER(ctypes.c_double)]),
    "ovx_server_run_trace": (_I, [_P, ctypes.c_void_p, _L, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    "ovx_server_destroy": (_I, [_P]),
    "ovx_corr_reference_create": (_I, [_I, _L, _L, _P, ctypes.POINTER(_I)]),
    "ovx_corr_reference_destroy": (_I, [_I]),
    "ovx_corr_plan_create": (_I, [_PP, _I, _L, _L, _U]),
    "ovx_corr_output_size": (_L, [_I, _L, _L]),
    "ovx_corr_exec": (_I, [_P, _I, _P, _P, _P, _P]),
    "ovx_conv_plan_create": (_I, [_PP, _I, _I, _P, _L, _L, _L, _U]),
    "ovx_conv_output_size": (_L, [_I, _L, _L, _L]),
    "ovx_conv_exec": (_I, [_P, _P, _P, _P]),
    "ovx_plan_destroy": (_I, [_P]),
    "ovx_plan_name": (ctypes.c_char_p, [_P]),
    "ovx_plan_device_bytes": (ctypes.c_size_t, [_P]),
}


class ServerConfig(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_int), ("max_jobs", ctypes.c_int), ("max_tiles", ctypes.c_long),
                ("hbm_budget_mb", ctypes.c_long), ("workspace_mb", ctypes.c_long), ("plain", ctypes.c_int)]


class Request(ctypes.Structure):
    _fields_ = [("t_arrival", ctypes.c_double), ("kind", ctypes.c_int), ("n", ctypes.c_long),
                ("rows", ctypes.c_long), ("x", ctypes.c_void_p), ("ldx", ctypes.c_long), ("y", ctypes.c_void_p),
                ("ldy", ctypes.c_long), ("ref", ctypes.c_int), ("bias", ctypes.c_int)]


REQ_FASTCONV = 1
REQ_FFTM = 2
REQ_CORR = 3


class OvxLibraryMissing(RuntimeError):
    pass


class OvxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ovx status {code}: {msg}")
        self.code = code


class OvxUnsupported(OvxError):
    """The request is valid but not handled by this backend (the reference's
    dispatcher would fall through to the next backend)."""


class OvxInvalid(OvxError, ValueError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OvxLibraryMissing(
                f"{LIB_PATH} is not built; run __graft_entry__.build() (make -C openvsip_amd/csrc)")
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def check(status):
    if status == OVX_OK:
        return
    msg = lib().ovx_last_error().decode(errors="replace")
    if status == OVX_ERR_UNSUPPORTED:
        raise OvxUnsupported(status, msg)
    if status == OVX_ERR_INVALID:
        raise OvxInvalid(status, msg)
    if status == OVX_ERR_NOMEM:
        raise MemoryError(msg)
    raise OvxError(status, msg)
