"""CPU-side checks of the C-ABI library: it loads, exports every symbol the
public header declares, and its pure host functions agree with the oracle.
No kernel is launched here."""
import ctypes
import os
import re

import pytest

import oracle as o
from openvsip_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "ovxhip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ovx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    # and the binding describes each of them
    assert set(syms) == set(_lib.SIGNATURES)


def test_output_sizes_match_oracle():
    lib = _lib.lib()
    for sup, code in _lib.SUPPORT.items():
        for m in (1, 3, 8):
            for n in (8, 9, 65536):
                assert lib.ovx_corr_output_size(code, m, n) == o.corr_output_size(sup, m, n)
                for d in (1, 2, 5):
                    assert lib.ovx_conv_output_size(code, m, n, d) == o.conv_output_size(sup, m, n, d)


def test_supported_lengths():
    lib = _lib.lib()
    for n in (2, 16, 2048, 3072, 12288, 16384, 10240, 9, 18, 9216, 65536, 262144):
        assert lib.ovx_supported_length(n) == 1, n
    for n in (7, 19, 68, 242, 252, 27, 16385):
        assert lib.ovx_supported_length(n) == 0, n
    assert lib.ovx_max_length() >= 16384


def test_invalid_arguments_return_status_not_crash():
    lib = _lib.lib()
    assert lib.ovx_fftm_plan_create(None, 4, 4, 0, -2, 1.0, 0) == _lib.OVX_ERR_INVALID
    assert b"null" in lib.ovx_last_error()
    h = ctypes.c_void_p()
    assert lib.ovx_fftm_plan_create(ctypes.byref(h), 4, 4, 5, -2, 1.0, 0) == _lib.OVX_ERR_INVALID
    assert lib.ovx_fftm_plan_create(ctypes.byref(h), 4, 4, 0, 3, 1.0, 0) == _lib.OVX_ERR_INVALID
    assert lib.ovx_corr_plan_create(ctypes.byref(h), 0, 10, 5, 0) == _lib.OVX_ERR_INVALID
    # real Fftm: even lengths only; axis and direction checked
    assert lib.ovx_fftm_real_plan_create(ctypes.byref(h), 4, 15, 0, -2, 1.0, 0) == _lib.OVX_ERR_UNSUPPORTED
    assert lib.ovx_fftm_real_plan_create(ctypes.byref(h), 4, 16, 2, -2, 1.0, 0) == _lib.OVX_ERR_INVALID
    assert lib.ovx_fftm_real_plan_create(ctypes.byref(h), 4, 16, 0, 0, 1.0, 0) == _lib.OVX_ERR_INVALID
    assert lib.ovx_fftm_real_out_of_place(None, None, 0, 0, None, 0, 0, 0, 0, None) == _lib.OVX_ERR_INVALID
    # the BF16x3 MFMA variant exists for n = 16384 on device pointers only
    assert lib.ovx_fastconv_plan_create(ctypes.byref(h), 4, 4096, 1.0, _lib.FLAG_BF16X3) == _lib.OVX_ERR_UNSUPPORTED
    assert lib.ovx_fastconv_plan_create(ctypes.byref(h), 4, 16384, 1.0,
                                        _lib.FLAG_BF16X3 | _lib.FLAG_HOST_POINTERS) == _lib.OVX_ERR_UNSUPPORTED
    assert lib.ovx_plan_destroy(None) == _lib.OVX_ERR_INVALID
    assert lib.ovx_fftm_out_of_place(None, None, 0, 0, None, 0, 0, 0, 0, None) == _lib.OVX_ERR_INVALID
    with pytest.raises(_lib.OvxInvalid):
        _lib.check(lib.ovx_plan_destroy(None))


def test_storage_host_side_needs_no_device():
    """User storage lives in the application's array: host access, release
    and admit involve no copy and no device (storage_manager's admit /
    release, src/ovxx/storage/manager.hpp:207-243)."""
    import numpy as np
    from openvsip_amd.api import Storage
    a = np.arange(16, dtype=np.complex64)
    s = Storage(a.nbytes, a)
    host = 1 << _lib.LOC_HOST
    assert s.ptr(Storage.HOST, Storage.READ).value == a.ctypes.data
    assert s.info() == (host, 0, 0)
    s.release()
    with pytest.raises(_lib.OvxError) as e:
        s.ptr(Storage.HOST, Storage.READ)
    assert e.value.code == _lib.OVX_ERR_STATE
    s.admit(update=True)
    assert s.ptr(Storage.HOST, Storage.WRITE).value == a.ctypes.data
    assert s.info() == (host, 0, 0)
    with pytest.raises(ValueError):
        Storage(a.nbytes + 8, a)
    lib = _lib.lib()
    assert lib.ovx_storage_create(None, 8, None) == _lib.OVX_ERR_INVALID
    assert lib.ovx_storage_ptr(None, 0, 1, None, None) == _lib.OVX_ERR_INVALID
    assert lib.ovx_storage_info(None, None, None, None) == _lib.OVX_ERR_INVALID
